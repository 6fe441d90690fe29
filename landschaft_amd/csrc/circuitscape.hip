// Circuitscape's raster modes on the graph engine, and the adjoint.
//
// Host side builds the node graph of a raster as the reference does
// (raster_graph, src/cpp/circuit/circuitscape.cpp:145-263: column-major node
// labels, short-circuit regions merged into one node, edges to the right,
// below and on both right-hand diagonals, duplicates summed, components);
// every solve, and the node currents, run on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <numeric>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../include/landschaft_circuit.h"
#include "common.hpp"
#include "host.hpp"
#include "system.hpp"

namespace lsc {

namespace {

constexpr double kShift = 10 * std::numeric_limits<double>::epsilon();  // CHOLMOD's, circuitscape.cpp:28
constexpr double kInf = std::numeric_limits<double>::infinity();

unsigned nblk(long long n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

// Raster graph (reference circuitscape.hpp:22-33).
struct RGraph {
    i64 nx = 0, ny = 0;
    i32 n = 0;
    std::vector<i32> nodemap;   // cell -> node
    std::vector<i64> node_cell; // node -> cell (single-cell nodes), -1 merged
    std::vector<i64> ptr;
    RawVec<i32> adj;
    RawVec<double> w;
    std::vector<double> degree;
    std::vector<i32> comp;
    i32 ncomp = 0;
};

// Circuitscape's cond_avg / weird_avg / res_avg / weirder_avg (circuitscape.cpp:31-34).
double average(double x, double y, bool diag, bool res) {
    if (res) return diag ? 1 / (1.4142135623730951 * (1 / x + 1 / y) / 2) : 1 / ((1 / x + 1 / y) / 2);
    return diag ? (x + y) / 2.8284271247461903 : (x + y) / 2;
}

// Components of the graph over edges of nonzero weight, numbered in the
// order of their smallest node: the numbering of a search started from every
// unlabelled node in turn. Union-find over host threads: a root is only ever
// linked under a smaller root, by a compare-and-swap on its own slot, so each
// tree's root is its component's smallest node whatever the interleaving, and
// finds halve their paths (every slot keeps pointing at a smaller node of its
// tree). The adjacency is symmetric, so each edge is taken once (t > v).
void label_components(RGraph& g) {
    const i32 n = g.n;
    g.comp.assign(n, -1);
    g.ncomp = 0;
    if (n == 0) return;
    std::unique_ptr<std::atomic<i32>[]> up(new std::atomic<i32>[n]);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 v = lo; v < hi; ++v) up[v].store(static_cast<i32>(v), std::memory_order_relaxed);
    });
    auto find = [&](i32 v) {
        for (;;) {
            const i32 p = up[v].load(std::memory_order_acquire);
            if (p == v) return v;
            const i32 gp = up[p].load(std::memory_order_acquire);
            if (gp != p) up[v].store(gp, std::memory_order_release);
            v = gp;
        }
    };
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 v = lo; v < hi; ++v)
            for (i64 k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
                i32 a = static_cast<i32>(v), b = g.adj[k];
                if (b <= a || g.w[k] == 0) continue;
                for (;;) {
                    a = find(a), b = find(b);
                    if (a == b) break;
                    if (a < b) std::swap(a, b);
                    i32 root = a;
                    if (up[a].compare_exchange_strong(root, b, std::memory_order_acq_rel)) break;
                }
            }
    });
    // roots (up[v] == v) in node order get the component numbers
    const i64 nt = host_chunk_count(n);
    std::vector<i32> base(static_cast<std::size_t>(nt) + 1, 0);
    host_chunks(n, [&](i64 t, i64 lo, i64 hi) {
        i32 c = 0;
        for (i64 v = lo; v < hi; ++v) c += up[v].load(std::memory_order_relaxed) == v;
        base[t + 1] = c;
    });
    for (i64 t = 0; t < nt; ++t) base[t + 1] += base[t];
    host_chunks(n, [&](i64 t, i64 lo, i64 hi) {
        i32 c = base[t];
        for (i64 v = lo; v < hi; ++v)
            if (up[v].load(std::memory_order_relaxed) == v) g.comp[v] = c++;
    });
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 v = lo; v < hi; ++v) {
            const i32 r = find(static_cast<i32>(v));  // roots' own slots are only read here
            if (r != v) g.comp[v] = g.comp[r];
        }
    });
    g.ncomp = base[nt];
}

// Without short-circuit regions every node is one cell and nodes are
// numbered in column-major cell order, so a node's neighbours in ascending
// node order are its cells' neighbours in column-major order: the CSR is
// written directly, with the same entries as the sorted edge list (no pair
// of cells is joined twice, and the edge average is symmetric in its two
// arguments, bit for bit).
// Column-major ranks (x outer, y inner) of the cells with gmap > 0 into
// rank[c] (-1 elsewhere), and each rank's cell into cell_of when given; the
// number of such cells. Two row-major passes over blocks of rows: each block
// counts its cells per column, then numbers them from the column's start
// plus the blocks above.
i32 column_ranks(const double* gmap, i64 nx, i64 ny, std::vector<i32>& rank, std::vector<i64>* cell_of) {
    const i64 n = nx * ny;
    rank.resize(n);
    const i64 nb = host_chunk_count(n);
    std::vector<std::vector<i32>> cnt(nb, std::vector<i32>(nx, 0));
    std::vector<std::thread> ts;
    auto rows = [&](i64 b) { return std::make_pair(ny * b / nb, ny * (b + 1) / nb); };
    for (i64 b = 0; b < nb; ++b)
        ts.emplace_back([&, b] {
            const auto [r0, r1] = rows(b);
            i32* c = cnt[b].data();
            for (i64 y = r0; y < r1; ++y)
                for (i64 x = 0; x < nx; ++x) c[x] += gmap[y * nx + x] > 0;
        });
    for (auto& t : ts) t.join();
    ts.clear();
    std::vector<i64> start(nx + 1, 0);
    for (i64 x = 0; x < nx; ++x) {
        i64 tot = 0;
        for (i64 b = 0; b < nb; ++b) tot += cnt[b][x];
        start[x + 1] = start[x] + tot;
    }
    const i64 total = start[nx];
    if (total >= (i64(1) << 31)) throw std::invalid_argument("too many nodes for 32-bit labels");
    if (cell_of) cell_of->resize(total);
    for (i64 b = 0; b < nb; ++b)
        ts.emplace_back([&, b] {
            const auto [r0, r1] = rows(b);
            std::vector<i64> next(nx);
            for (i64 x = 0; x < nx; ++x) {
                i64 o = start[x];
                for (i64 bb = 0; bb < b; ++bb) o += cnt[bb][x];
                next[x] = o;
            }
            for (i64 y = r0; y < r1; ++y)
                for (i64 x = 0; x < nx; ++x) {
                    const i64 c = y * nx + x;
                    if (gmap[c] > 0) {
                        const i32 v = static_cast<i32>(next[x]++);
                        rank[c] = v;
                        if (cell_of) (*cell_of)[v] = c;
                    } else {
                        rank[c] = -1;
                    }
                }
        });
    for (auto& t : ts) t.join();
    return static_cast<i32>(total);
}

RGraph raster_graph_cells(const double* gmap, i64 nx, i64 ny, bool avg_res, bool four) {
    RGraph g;
    g.nx = nx, g.ny = ny;
    g.n = column_ranks(gmap, nx, ny, g.nodemap, &g.node_cell);
    progress("raster graph: %d nodes labelled", g.n);
    const i32 cntn = g.n;
    static const int k8[8][2] = {{-1, -1}, {-1, 0}, {-1, 1}, {0, -1}, {0, 1}, {1, -1}, {1, 0}, {1, 1}};
    static const int k4[4][2] = {{-1, 0}, {0, -1}, {0, 1}, {1, 0}};
    const int nd = four ? 4 : 8;
    g.ptr.assign(static_cast<std::size_t>(cntn) + 1, 0);
    host_chunks(cntn, [&](i64, i64 lo, i64 hi) {
        for (i64 v = lo; v < hi; ++v) {
            const i64 c = g.node_cell[v], x = c % nx, y = c / nx;
            i64 deg = 0;
            for (int d = 0; d < nd; ++d) {
                const int dx = four ? k4[d][0] : k8[d][0], dy = four ? k4[d][1] : k8[d][1];
                const i64 xx = x + dx, yy = y + dy;
                if (xx >= 0 && xx < nx && yy >= 0 && yy < ny && g.nodemap[yy * nx + xx] >= 0) ++deg;
            }
            g.ptr[v + 1] = deg;
        }
    });
    host_prefix(cntn, g.ptr.data() + 1, g.ptr.data());
    g.adj.resize(g.ptr[cntn]);
    g.w.resize(g.ptr[cntn]);
    g.degree.assign(cntn, 0.0);
    host_chunks(cntn, [&](i64, i64 lo, i64 hi) {
        for (i64 v = lo; v < hi; ++v) {
            const i64 c = g.node_cell[v], x = c % nx, y = c / nx;
            i64 k = g.ptr[v];
            double dsum = 0.0;
            for (int d = 0; d < nd; ++d) {
                const int dx = four ? k4[d][0] : k8[d][0], dy = four ? k4[d][1] : k8[d][1];
                const i64 xx = x + dx, yy = y + dy;
                if (xx < 0 || xx >= nx || yy < 0 || yy >= ny) continue;
                const i64 e = yy * nx + xx;
                const i32 u = g.nodemap[e];
                if (u < 0) continue;
                const double w = average(gmap[c], gmap[e], dx != 0 && dy != 0, avg_res);
                g.adj[k] = u;
                g.w[k] = w;
                dsum += w;
                ++k;
            }
            g.degree[v] = dsum;
        }
    });
    progress("raster graph: %lld adjacency entries", static_cast<long long>(g.ptr[cntn]));
    label_components(g);
    progress("raster graph: %d components", g.ncomp);
    return g;
}

// The graph with short-circuit regions (polygon ids != 0): every region with
// at least one cell of gmap > 0 is one node, numbered at its first such cell
// in column-major order, and takes all of its cells, NoData ones too; the
// other nodes keep their cells' column-major order. A node's entries are its
// cells' branches (right, down and the two right diagonals of every emitting
// cell, in column-major emitter order), self loops left out, stably sorted by
// neighbour, repeats summed in that order: per node that is a sort by
// (neighbour, emitter position, direction), so nodes are built independently
// over host threads.
RGraph raster_graph_regions(const double* gmap, const std::int64_t* poly, i64 nx, i64 ny, bool avg_res,
                            bool four) {
    RGraph g;
    g.nx = nx, g.ny = ny;
    const i64 n = nx * ny;
    std::vector<i32> rank;
    std::vector<i64> rank_cell;
    const i32 nvalid = column_ranks(gmap, nx, ny, rank, &rank_cell);
    // region cells, row-major
    std::vector<i64> rcells;
    {
        const i64 nt = host_chunk_count(n);
        std::vector<std::vector<i64>> part(static_cast<std::size_t>(nt));
        host_chunks(n, [&](i64 t, i64 lo, i64 hi) {
            for (i64 c = lo; c < hi; ++c)
                if (poly[c] != 0) part[t].push_back(c);
        });
        for (auto& q : part) rcells.insert(rcells.end(), q.begin(), q.end());
    }
    // each region's first valid cell (smallest rank); its other valid cells
    // are merged into it
    std::unordered_map<std::int64_t, i32> first;
    for (i64 c : rcells)
        if (rank[c] >= 0) {
            auto it = first.emplace(poly[c], rank[c]).first;
            it->second = std::min(it->second, rank[c]);
        }
    std::vector<std::uint8_t> keep(static_cast<std::size_t>(nvalid), 1);
    for (i64 c : rcells)
        if (rank[c] >= 0 && rank[c] != first[poly[c]]) keep[rank[c]] = 0;
    std::vector<i32> id(static_cast<std::size_t>(nvalid) + 1);
    host_prefix(nvalid, keep.data(), id.data());
    g.n = id[nvalid];
    // ranks to nodes; region cells to their region's node
    g.nodemap.resize(n);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 c = lo; c < hi; ++c) g.nodemap[c] = rank[c] >= 0 ? id[rank[c]] : -1;
    });
    std::vector<i32> cells_per(static_cast<std::size_t>(g.n), 1);
    std::unordered_map<i32, std::vector<i64>> region_cells;  // nodes of more than one cell
    {
        for (i64 c : rcells) {
            auto it = first.find(poly[c]);
            if (it == first.end()) continue;
            const i32 v = id[it->second];
            g.nodemap[c] = v;
            region_cells[v].push_back(c);
        }
        for (auto it = region_cells.begin(); it != region_cells.end();) {
            if (it->second.size() <= 1) {
                it = region_cells.erase(it);
            } else {
                cells_per[it->first] = static_cast<i32>(it->second.size());
                ++it;
            }
        }
    }
    g.node_cell.resize(g.n);
    host_chunks(nvalid, [&](i64, i64 lo, i64 hi) {
        for (i64 r = lo; r < hi; ++r)
            if (keep[r]) g.node_cell[id[r]] = cells_per[id[r]] > 1 ? -1 : rank_cell[r];
    });
    std::vector<i32>().swap(rank);
    struct Entry {
        i32 nb;
        i64 key;
        double w;
    };
    // the entries of node v (cells: its cells) appended to out, unsorted
    auto gather = [&](i32 v, const i64* cells, std::size_t count, std::vector<Entry>& out) {
        for (std::size_t i = 0; i < count; ++i) {
            const i64 c = cells[i], x = c % nx, y = c / nx;
            const auto fwd = [&](i64 d, int dir, bool diag) {
                const i32 u = g.nodemap[d];
                if (u >= 0 && u != v) out.push_back({u, (x * ny + y) * 4 + dir, average(gmap[c], gmap[d], diag, avg_res)});
            };
            const auto back = [&](i64 e, i64 ex, i64 ey, int dir, bool diag) {
                const i32 u = g.nodemap[e];
                if (u >= 0 && u != v) out.push_back({u, (ex * ny + ey) * 4 + dir, average(gmap[e], gmap[c], diag, avg_res)});
            };
            if (x + 1 < nx) fwd(c + 1, 0, false);
            if (y + 1 < ny) fwd(c + nx, 1, false);
            if (!four && y + 1 < ny && x + 1 < nx) fwd(c + nx + 1, 2, true);
            if (!four && y > 0 && x + 1 < nx) fwd(c - nx + 1, 3, true);
            if (x > 0) back(c - 1, x - 1, y, 0, false);
            if (y > 0) back(c - nx, x, y - 1, 1, false);
            if (!four && x > 0 && y > 0) back(c - nx - 1, x - 1, y - 1, 2, true);
            if (!four && x > 0 && y + 1 < ny) back(c + nx - 1, x - 1, y + 1, 3, true);
        }
    };
    // node v's row: entries by (neighbour, emission), repeats summed in order
    auto row = [&](i32 v, std::vector<Entry>& ent, auto&& put) {
        ent.clear();
        auto rc = region_cells.find(v);
        if (rc != region_cells.end()) gather(v, rc->second.data(), rc->second.size(), ent);
        else if (g.node_cell[v] >= 0) gather(v, &g.node_cell[v], 1, ent);
        std::sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) {
            return a.nb != b.nb ? a.nb < b.nb : a.key < b.key;
        });
        for (std::size_t k = 0; k < ent.size();) {
            double w = ent[k].w;
            std::size_t e = k + 1;
            for (; e < ent.size() && ent[e].nb == ent[k].nb; ++e) w += ent[e].w;
            put(ent[k].nb, w);
            k = e;
        }
    };
    const i32 m = g.n;
    g.ptr.assign(static_cast<std::size_t>(m) + 1, 0);
    host_chunks(m, [&](i64, i64 lo, i64 hi) {
        std::vector<Entry> ent;
        for (i64 v = lo; v < hi; ++v) {
            i64 k = 0;
            row(static_cast<i32>(v), ent, [&](i32, double) { ++k; });
            g.ptr[v + 1] = k;
        }
    });
    host_prefix(m, g.ptr.data() + 1, g.ptr.data());
    g.adj.resize(static_cast<std::size_t>(g.ptr[m]));
    g.w.resize(static_cast<std::size_t>(g.ptr[m]));
    g.degree.assign(m, 0.0);
    host_chunks(m, [&](i64, i64 lo, i64 hi) {
        std::vector<Entry> ent;
        for (i64 v = lo; v < hi; ++v) {
            i64 k = g.ptr[v];
            double d = 0.0;
            row(static_cast<i32>(v), ent, [&](i32 u, double w) {
                g.adj[k] = u;
                g.w[k++] = w;
                d += w;
            });
            g.degree[v] = d;
        }
    });
    progress("raster graph: %d nodes, %lld adjacency entries (regions)", m, static_cast<long long>(g.ptr[m]));
    label_components(g);
    return g;
}

RGraph raster_graph(const double* gmap, const std::int64_t* poly, i64 nx, i64 ny, bool avg_res, bool four) {
    const i64 n = nx * ny;
    std::atomic<bool> any_poly{false};
    if (poly)
        host_chunks(n, [&](i64, i64 lo, i64 hi) {
            for (i64 c = lo; c < hi && !any_poly.load(std::memory_order_relaxed); ++c)
                if (poly[c] != 0) any_poly.store(true, std::memory_order_relaxed);
        });
    return any_poly ? raster_graph_regions(gmap, poly, nx, ny, avg_res, four)
                    : raster_graph_cells(gmap, nx, ny, avg_res, four);
}

// A network's graph (reference network_graph, circuitscape.cpp:265-338): the
// edge list as a sparse matrix whose repeated positions add up in input order,
// made symmetric (a position holds its value plus its mirror's), self loops
// and zero sums left out; each node's neighbours ascending. No grid: nx = ny =
// 0 and no node has a cell.
RGraph network_graph(const i32* a, const i32* b, const double* w, i64 ne, i32 m) {
    for (i64 k = 0; k < ne; ++k)
        if (a[k] < 0 || b[k] < 0 || a[k] >= m || b[k] >= m) throw std::invalid_argument("node index out of range");
    // 1. positions (column-major keys); a stable order keeps repeats in input order
    const auto key = [&](i64 k) { return static_cast<std::uint64_t>(b[k]) * static_cast<std::uint64_t>(m) + a[k]; };
    std::vector<i64> order(static_cast<std::size_t>(ne));
    std::iota(order.begin(), order.end(), i64{0});
    std::stable_sort(order.begin(), order.end(), [&](i64 x, i64 y) { return key(x) < key(y); });
    std::vector<i32> pr, pc;  // the distinct off-diagonal positions and their summed values
    std::vector<double> pv;
    for (std::size_t q = 0; q < order.size();) {
        const i64 k0 = order[q];
        double sum = w[k0];
        std::size_t t = q + 1;
        for (; t < order.size() && key(order[t]) == key(k0); ++t) sum += w[order[t]];
        if (a[k0] != b[k0]) pr.push_back(a[k0]), pc.push_back(b[k0]), pv.push_back(sum);
        q = t;
    }
    // 2. every position and its mirror under the column of each end: at most two
    // values per (node, neighbour), added (the sum of two does not depend on order)
    RGraph g;
    g.n = m;
    g.node_cell.assign(m, -1);
    std::vector<i64> slot(static_cast<std::size_t>(m) + 1, 0);
    for (std::size_t e = 0; e < pv.size(); ++e) ++slot[pc[e] + 1], ++slot[pr[e] + 1];
    for (i32 v = 0; v < m; ++v) slot[v + 1] += slot[v];
    std::vector<std::pair<i32, double>> nb(static_cast<std::size_t>(slot[m]));
    {
        std::vector<i64> fill(slot.begin(), slot.end() - 1);
        for (std::size_t e = 0; e < pv.size(); ++e) {
            nb[fill[pc[e]]++] = {pr[e], pv[e]};
            nb[fill[pr[e]]++] = {pc[e], pv[e]};
        }
    }
    g.ptr.assign(static_cast<std::size_t>(m) + 1, 0);
    g.degree.assign(m, 0.0);
    for (i32 v = 0; v < m; ++v) {
        auto first = nb.begin() + slot[v], last = nb.begin() + slot[v + 1];
        std::sort(first, last, [](const auto& x, const auto& y) { return x.first < y.first; });
        double deg = 0.0;
        for (auto it = first; it != last;) {
            double sum = it->second;
            auto next = it + 1;
            if (next != last && next->first == it->first) sum += (next++)->second;
            if (sum != 0) {
                g.adj.push_back(it->first);
                g.w.push_back(sum);
                deg += sum;
            }
            it = next;
        }
        g.degree[v] = deg;
        g.ptr[v + 1] = static_cast<i64>(g.adj.size());
    }
    label_components(g);
    return g;
}

struct Members {
    std::vector<i64> start;
    std::vector<i32> mem;
};
Members members_of(const RGraph& g) {
    Members m;
    m.start.assign(g.ncomp + 1, 0);
    for (i32 v = 0; v < g.n; ++v) ++m.start[g.comp[v] + 1];
    for (i32 k = 0; k < g.ncomp; ++k) m.start[k + 1] += m.start[k];
    m.mem.resize(g.n);
    std::vector<i64> fill(m.start.begin(), m.start.end() - 1);
    for (i32 v = 0; v < g.n; ++v) m.mem[fill[g.comp[v]]++] = v;
    return m;
}

// The grounded Laplacian of one component without the dropped nodes, plus
// `finite` on the diagonal and the shift (reference factor_component,
// circuitscape.cpp:92-141). `nodes` receives the caller's node of every
// unknown.
template <class Drop>
HostMatrix component_matrix(const RGraph& g, const i32* mem, i64 count, const double* finite, Drop drop,
                            std::vector<i32>& nodes) {
    HostMatrix M;
    // node -> unknown, -1 elsewhere; kept between calls and reset after use
    // (bound to a reference: the host threads below use this thread's copy)
    static thread_local std::vector<i32> local_of;
    std::vector<i32>& local = local_of;
    if (local.size() < static_cast<std::size_t>(g.n)) local.assign(g.n, -1);
    // unknowns: the kept members in member order, numbered from per-chunk counts
    const i64 nt = host_chunk_count(count);
    std::vector<i64> kept(static_cast<std::size_t>(nt) + 1, 0);
    std::vector<std::array<i64, 4>> box(static_cast<std::size_t>(nt), {g.nx, -1, g.ny, -1});
    host_chunks(count, [&](i64 t, i64 lo, i64 hi) {
        i64 c = 0;
        for (i64 p = lo; p < hi; ++p) c += !drop(mem[p]);
        kept[t + 1] = c;
    });
    for (i64 t = 0; t < nt; ++t) kept[t + 1] += kept[t];
    nodes.resize(static_cast<std::size_t>(kept[nt]));
    host_chunks(count, [&](i64 t, i64 lo, i64 hi) {
        i64 i = kept[t];
        auto& b = box[t];
        for (i64 p = lo; p < hi; ++p) {
            const i32 v = mem[p];
            if (drop(v)) continue;
            local[v] = static_cast<i32>(i);
            nodes[i++] = v;
            const i64 c = g.node_cell[v];
            if (c >= 0) {
                b[0] = std::min(b[0], c % g.nx), b[1] = std::max(b[1], c % g.nx);
                b[2] = std::min(b[2], c / g.nx), b[3] = std::max(b[3], c / g.nx);
            }
        }
    });
    i64 x0 = g.nx, x1 = -1, y0 = g.ny, y1 = -1;
    for (const auto& b : box) x0 = std::min(x0, b[0]), x1 = std::max(x1, b[1]), y0 = std::min(y0, b[2]), y1 = std::max(y1, b[3]);
    const i64 n = static_cast<i64>(nodes.size());
    M.n = static_cast<int>(n);
    M.bx = x1 >= x0 ? x1 - x0 + 1 : 1;
    M.by = y1 >= y0 ? y1 - y0 + 1 : 1;
    M.cell.resize(n);
    M.ptr.assign(n + 1, 0);
    M.diag.resize(n);
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 i = lo; i < hi; ++i) {
            const i32 v = nodes[i];
            const i64 c = g.node_cell[v];
            M.cell[i] = c >= 0 ? (c / g.nx - y0) * M.bx + (c % g.nx - x0) : -1;
            i64 k = 0;
            for (i64 e = g.ptr[v]; e < g.ptr[v + 1]; ++e) k += local[g.adj[e]] >= 0;
            M.ptr[i + 1] = k;
            M.diag[i] = (finite ? g.degree[v] + finite[v] : g.degree[v]) + kShift;
        }
    });
    host_prefix(n, M.ptr.data() + 1, M.ptr.data());
    M.adj.resize(static_cast<std::size_t>(M.ptr[n]));
    M.off.resize(static_cast<std::size_t>(M.ptr[n]));
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 i = lo; i < hi; ++i) {
            const i32 v = nodes[i];
            i64 k = M.ptr[i];
            for (i64 e = g.ptr[v]; e < g.ptr[v + 1]; ++e) {
                const i32 u = local[g.adj[e]];
                if (u < 0) continue;
                M.adj[k] = u;
                M.off[k++] = -g.w[e];
            }
        }
    });
    host_chunks(n, [&](i64, i64 lo, i64 hi) {
        for (i64 i = lo; i < hi; ++i) local[nodes[i]] = -1;
    });
    return M;
}

// ---- node currents on the GPU (reference graph_node_currents,
// circuitscape.cpp:38-73; on a column-major raster graph it is also the
// grid's node_currents, grid.cpp:19-71: an edge's current is computed from
// either end with the opposite sign, exactly).

struct DevGraph {
    DevBuf<long long> ptr;
    DevBuf<int> adj;
    DevBuf<double> w;
    void upload(const RGraph& g, hipStream_t s) {
        std::vector<long long> p(g.ptr.begin(), g.ptr.end());
        ptr = to_device(p, s);
        adj = to_device(g.adj, s);
        w = g.w.empty() ? to_device(std::vector<double>{0.0}, s) : to_device(g.w, s);
    }
};

__global__ void k_gc_max(int cnt, const int* __restrict__ mem, const long long* __restrict__ ptr,
                         const int* __restrict__ adj, const double* __restrict__ w, const double* __restrict__ v,
                         double* __restrict__ partial) {
    __shared__ double red[2][4];
    double mi = -INFINITY, mo = -INFINITY;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < cnt) {
        const int c = mem[p];
        for (long long e = ptr[c]; e < ptr[c + 1]; ++e) {
            const int r = adj[e];
            const double b = fabs(w[e]) * (v[r] - v[c]);
            if (r < c) mi = fmax(mi, b);
            else mo = fmax(mo, b);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mi = fmax(mi, __shfl_xor(mi, o));
        mo = fmax(mo, __shfl_xor(mo, o));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[0][wave] = mi, red[1][wave] = mo;
    __syncthreads();
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
        partial[2 * blockIdx.x + 1] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
    }
}

__global__ void k_gc_reduce(int nb, const double* __restrict__ partial, double* __restrict__ mx) {
    __shared__ double red[2][4];
    double mi = -INFINITY, mo = -INFINITY;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) mi = fmax(mi, partial[2 * i]), mo = fmax(mo, partial[2 * i + 1]);
    for (int o = 32; o > 0; o >>= 1) {
        mi = fmax(mi, __shfl_xor(mi, o));
        mo = fmax(mo, __shfl_xor(mo, o));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[0][wave] = mi, red[1][wave] = mo;
    __syncthreads();
    if (threadIdx.x == 0) {
        mx[0] = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
        mx[1] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
    }
}

__global__ void k_gc_cur(int cnt, const int* __restrict__ mem, const long long* __restrict__ ptr,
                         const int* __restrict__ adj, const double* __restrict__ w, const double* __restrict__ v,
                         const double* __restrict__ finite, const double* __restrict__ mx, double* __restrict__ cur) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cnt) return;
    const double maxin = mx[0], maxout = mx[1];
    const int c = mem[p];
    double in = 0.0, out = 0.0;
    for (long long e = ptr[c]; e < ptr[c + 1]; ++e) {
        const int r = adj[e];
        const double b = fabs(w[e]) * (v[r] - v[c]);
        if (b > 0) {
            if (!(fabs(b / maxin) < 1e-8)) in += b;
        } else if (b < 0) {
            if (!(fabs(b / maxout) < 1e-8)) out -= b;
        }
    }
    if (finite) {
        const double f = finite[c] * v[c];
        if (f < 0) in -= f;
        else if (f > 0) out += f;
    }
    cur[c] = in > out ? in : out;
}

// Node currents of one component (`d_mem`: its nodes, ascending) from
// voltages `d_v` (by node) into `d_cur` (by node).
void graph_currents(const DevGraph& G, const int* d_mem, int cnt, const double* d_v, const double* d_finite,
                    double* d_cur, DevBuf<double>& scratch, hipStream_t s) {
    if (cnt <= 0) return;
    const unsigned nb = nblk(cnt, 256);
    if (scratch.n < 2 * nb + 2) scratch.alloc(2 * nb + 2);
    hipLaunchKernelGGL(k_gc_max, dim3(nb), dim3(256), 0, s, cnt, d_mem, G.ptr.p, G.adj.p, G.w.p, d_v, scratch.p);
    hipLaunchKernelGGL(k_gc_reduce, dim3(1), dim3(256), 0, s, static_cast<int>(nb), scratch.p, scratch.p + 2 * nb);
    hipLaunchKernelGGL(k_gc_cur, dim3(nb), dim3(256), 0, s, cnt, d_mem, G.ptr.p, G.adj.p, G.w.p, d_v, d_finite,
                       scratch.p + 2 * nb, d_cur);
}

// One advanced problem: the nodes with a nonzero source or ground (ascending)
// and their per-node sums.
struct Problem {
    std::vector<i32> node;
    std::vector<double> src, gnd;
};

// From dense per-node arrays.
Problem problem_from_dense(const std::vector<double>& src, const std::vector<double>& gnd) {
    Problem p;
    for (std::size_t v = 0; v < src.size(); ++v)
        if (src[v] != 0 || gnd[v] != 0) {
            p.node.push_back(static_cast<i32>(v));
            p.src.push_back(src[v]);
            p.gnd.push_back(gnd[v]);
        }
    return p;
}

// The problem's values after resolve_conflicts (reference
// circuitscape.cpp:881-900; 0 keepall, 1 rmvsrc, 2 rmvgnd, 3 rmvall): finite
// grounds stay on the diagonal as given (before the policy), the policy on
// nodes that are both, and a direct ground under a positive source dropped.
// Nodes whose source and ground are both zero play no part, so the sparse
// form is exact.
struct Resolved {
    std::vector<i32> node;
    std::vector<double> src, gnd, finite;
    bool any_finite = false;
    double at(const std::vector<double>& a, i32 v) const {
        auto it = std::lower_bound(node.begin(), node.end(), v);
        return it != node.end() && *it == v ? a[it - node.begin()] : 0.0;
    }
    std::vector<double> dense(const std::vector<double>& a, i32 n) const {
        std::vector<double> d(n, 0.0);
        for (std::size_t i = 0; i < node.size(); ++i) d[node[i]] = a[i];
        return d;
    }
};

Resolved resolve(const Problem& pb, int policy) {
    Resolved r;
    r.node = pb.node;
    r.src = pb.src;
    r.gnd = pb.gnd;
    r.finite.assign(pb.node.size(), 0.0);
    for (std::size_t i = 0; i < r.node.size(); ++i) {
        r.finite[i] = r.gnd[i] < kInf ? r.gnd[i] : 0.0;
        if (r.finite[i] != 0) r.any_finite = true;
    }
    for (std::size_t i = 0; i < r.node.size(); ++i) {
        if (!(r.src[i] != 0 && r.gnd[i] != 0)) continue;
        if (policy == 1 || policy == 3) r.src[i] = 0.0;
        else if (policy == 2) r.gnd[i] = 0.0;
    }
    for (std::size_t i = 0; i < r.node.size(); ++i)
        if (r.gnd[i] == kInf && r.src[i] > 0) r.gnd[i] = 0.0;
    return r;
}

struct NodeSol {
    std::vector<double> voltage, current;
    std::vector<i32> active;
    bool solved = false;
};

struct Ctx {
    int precision;
    double tol;
    int maxit;
    hipStream_t s;
    IrStats st;
};

// A solver's host analysis, factor and solve times into the call's totals.
void tally(Ctx& cx, const SystemSolver& S) {
    cx.st.ms_fill += S.analyze_ms;
    cx.st.ms_factor += S.factor_ms;
    cx.st.ms_solve += S.solve_ms;
    cx.st.fallback += S.fallbacks();
}

// The precision of a factorisation that serves `nrhs` right-hand sides.
// LSC_PRECISION_AUTO: FP64 from kAutoFp64Rhs of them, mixed below. On these
// graphs an FP32 factor preconditions about seven CG steps per right-hand
// side (each a forward and back sweep and a residual) where an FP64 factor,
// about 2.7 times the FP32 one's cost, solves in one sweep pair: at 8192^2
// the two break even near 7.5 right-hand sides, at 2048^2 near 2.5.
constexpr long long kAutoFp64Rhs = 8;
int solver_precision(int p, long long nrhs) {
    if (p != LSC_PRECISION_AUTO) return p;
    return nrhs >= kAutoFp64Rhs ? LSC_PRECISION_FP64 : LSC_PRECISION_MIXED;
}

Ctx make_ctx(const lsc_solve_options* o) {
    Ctx c{};
    c.precision = o ? o->precision : LSC_PRECISION_MIXED;
    c.tol = o && o->tol > 0 ? o->tol : 1e-12;
    c.maxit = o && o->max_refine > 0 ? o->max_refine : 30;
    c.s = o ? reinterpret_cast<hipStream_t>(o->stream) : nullptr;
    return c;
}

// corr(a, b) = sum over the neighbours e of special node a outside the
// specials of w_ae Z_b[e], so that S = A_FF - A_FU A_UU^-1 A_UF = A_FF + corr
// (A_UF = -w, Z = A_UU^-1 A_UF): one thread per (a, b).
__global__ void k_schur(int nf, long long na, const long long* __restrict__ sptr, const int* __restrict__ sidx,
                        const double* __restrict__ sw, const double* __restrict__ Z, double* __restrict__ S) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(nf) * nf) return;
    const int a = static_cast<int>(t / nf), b = static_cast<int>(t - static_cast<long long>(a) * nf);
    const double* Zb = Z + static_cast<long long>(b) * na;
    double acc = 0.0;
    for (long long k = sptr[a]; k < sptr[a + 1]; ++k) acc += sw[k] * Zb[sidx[k]];
    S[t] = acc;
}

// B = A_UF, column f of nf: entry (f, e) = -w (Z holds B before the solves).
__global__ void k_schur_b(long long cnt, long long na, const int* __restrict__ bf, const int* __restrict__ be,
                          const double* __restrict__ bw, double* __restrict__ Z) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < cnt) Z[static_cast<long long>(bf[t]) * na + be[t]] = -bw[t];
}

__global__ void k_fill_d(double* __restrict__ p, long long n, double v) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Voltages of the non-special nodes of one problem: v[nodes[e]] = -sum_f Z_f[e] y_f,
// and the specials' own values.
__global__ void k_schur_v(long long na, int nf, const double* __restrict__ Z, const double* __restrict__ yf,
                          const int* __restrict__ nodes, double* __restrict__ v) {
    const long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= na) return;
    double t = 0.0;
    for (int f = 0; f < nf; ++f) {
        const double y = yf[f];
        if (y != 0.0) t += Z[static_cast<long long>(f) * na + e] * y;
    }
    v[nodes[e]] = -t;
}

__global__ void k_scatter_d(int cnt, const int* __restrict__ idx, const double* __restrict__ val,
                            double* __restrict__ v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cnt) v[idx[t]] = val[t];
}

// Dense SPD solve A y = y in place (A row-major nu x nu, lower Cholesky).
void dense_spd_solve(std::vector<double>& A, std::vector<double>& y, std::size_t nu) {
    for (std::size_t j = 0; j < nu; ++j) {
        double d = A[j * nu + j];
        for (std::size_t q = 0; q < j; ++q) d -= A[j * nu + q] * A[j * nu + q];
        if (!(d > 0)) throw std::runtime_error("a circuit component is not positive definite");
        d = std::sqrt(d);
        A[j * nu + j] = d;
        for (std::size_t i = j + 1; i < nu; ++i) {
            double t = A[i * nu + j];
            for (std::size_t q = 0; q < j; ++q) t -= A[i * nu + q] * A[j * nu + q];
            A[i * nu + j] = t / d;
        }
    }
    for (std::size_t i = 0; i < nu; ++i) {
        double t = y[i];
        for (std::size_t q = 0; q < i; ++q) t -= A[i * nu + q] * y[q];
        y[i] = t / A[i * nu + i];
    }
    for (std::size_t i = nu; i-- > 0;) {
        double t = y[i];
        for (std::size_t q = i + 1; q < nu; ++q) t -= A[q * nu + i] * y[q];
        y[i] = t / A[i * nu + i];
    }
}

// Receives every solved (problem, component): the full per-node voltage and
// current vectors on the device (valid on the component's members).
struct Sink {
    virtual ~Sink() = default;
    virtual void component(std::size_t p, i32 k, const Resolved& r, const double* d_v, const double* d_cur) = 0;
};

// Advanced-mode problems on one graph (the reference's raster_advanced and
// raster_advanced_many, circuitscape.cpp:393-468, 904-1083). Per component:
// with two or more problems and at most 2048 special nodes (sources or
// grounds of any problem), one factorisation without the specials and the
// Schur complement on them, S = A_FF - A_FU A_UU^-1 A_UF, built from nf
// solves; each problem then solves its free specials densely and gets the
// rest as -Z y_F. Otherwise every problem is factorised on its own.
// Components go in order, and within one its problems in order. Returns
// whether each problem had a component solved.
std::vector<char> advanced_many(const RGraph& g, const DevGraph& G, const std::vector<Problem>& probs,
                                const std::vector<i32>& only, int policy, bool allow_schur, Ctx& cx, Sink& sink) {
    const std::size_t np = probs.size();
    const i32 n = g.n;
    std::vector<Resolved> res(np);
    std::vector<std::uint8_t> special(n, 0);
    for (std::size_t p = 0; p < np; ++p) {
        res[p] = resolve(probs[p], policy);
        for (std::size_t i = 0; i < res[p].node.size(); ++i)
            if (res[p].src[i] != 0 || res[p].gnd[i] != 0 || res[p].finite[i] != 0) special[res[p].node[i]] = 1;
    }
    const Members M = members_of(g);
    // the problems to solve in every component (advanced_kernel's tests):
    // source and ground sums over the component's nodes, in node order
    std::vector<std::vector<std::size_t>> todo(g.ncomp);
    std::vector<char> solved(np, 0);
    for (std::size_t p = 0; p < np; ++p) {
        std::vector<double> ss(g.ncomp, 0.0), sg(g.ncomp, 0.0);
        const Resolved& r = res[p];
        for (std::size_t i = 0; i < r.node.size(); ++i) {
            const i32 k = g.comp[r.node[i]];
            ss[k] += r.src[i];
            sg[k] += r.gnd[i];
        }
        for (i32 k = 0; k < g.ncomp; ++k) {
            if (only[p] == -2 || (only[p] >= 0 && g.comp[only[p]] != k)) continue;
            if (ss[k] != 0 && sg[k] != 0) {
                todo[k].push_back(p);
                solved[p] = 1;
            }
        }
    }
    DevBuf<double> dv(n ? n : 1), dcur(n ? n : 1), dfin(n ? n : 1), scratch;
    DevBuf<int> dmem = to_device(M.mem.empty() ? std::vector<i32>{0} : M.mem, cx.s);
    std::vector<i32> nodes;
    // currents of problem p on component k from the voltages in dv
    auto finish = [&](std::size_t p, i32 k) {
        const Resolved& r = res[p];
        const i64 count = M.start[k + 1] - M.start[k];
        if (r.any_finite) {
            const std::vector<double> fin = r.dense(r.finite, n);
            dfin.upload(fin.data(), n, cx.s);
        }
        graph_currents(G, dmem.p + M.start[k], static_cast<int>(count), dv.p, r.any_finite ? dfin.p : nullptr, dcur.p,
                       scratch, cx.s);
        sink.component(p, k, r, dv.p, dcur.p);
    };
    for (i32 k = 0; k < g.ncomp; ++k) {
        const auto& pk = todo[k];
        if (pk.empty()) continue;
        const i32* mem = M.mem.data() + M.start[k];
        const i64 count = M.start[k + 1] - M.start[k];
        std::vector<i32> F;
        for (i64 q = 0; q < count; ++q)
            if (special[mem[q]]) F.push_back(mem[q]);
        const std::size_t nf = F.size();
        if (!allow_schur || nf > 2048 || pk.size() < 2) {
            for (std::size_t p : pk) {
                const Resolved& r = res[p];
                const std::vector<double> fin = r.any_finite ? r.dense(r.finite, n) : std::vector<double>{};
                HostMatrix A = component_matrix(g, mem, count, r.any_finite ? fin.data() : nullptr,
                                                [&](i32 v) { return r.at(r.gnd, v) == kInf; }, nodes);
                std::vector<double> full(n, 0.0);
                if (A.n > 0) {
                    SystemSolver S(solver_precision(cx.precision, 1), cx.s);
                    S.factor(A);
                    std::vector<double> b(A.n), xv(A.n);
                    for (int i = 0; i < A.n; ++i) b[i] = r.at(r.src, nodes[i]);
                    S.solve(1, b.data(), xv.data(), cx.tol, cx.maxit, cx.st);
                    tally(cx, S);
                    for (int i = 0; i < A.n; ++i) full[nodes[i]] = xv[i];
                }
                dv.upload(full.data(), n, cx.s);
                finish(p, k);
            }
            continue;
        }
        progress("component %d: %lld nodes, %zu special, %zu problems (Schur)", k, static_cast<long long>(count), nf,
                 pk.size());
        HostMatrix A = component_matrix(g, mem, count, nullptr, [&](i32 v) { return special[v] != 0; }, nodes);
        const long long na = A.n;
        std::vector<i32> uidx(n, -1), fidx(n, -1);
        for (long long e = 0; e < na; ++e) uidx[nodes[e]] = static_cast<i32>(e);
        for (std::size_t f = 0; f < nf; ++f) fidx[F[f]] = static_cast<i32>(f);
        // the couplings of every special to the rest, and S's special-special part
        std::vector<long long> sptr(nf + 1, 0);
        std::vector<int> sidx, bf;
        std::vector<double> sw;
        std::vector<double> S(nf * nf, 0.0);
        for (std::size_t f = 0; f < nf; ++f) {
            S[f * nf + f] = g.degree[F[f]];
            for (i64 e = g.ptr[F[f]]; e < g.ptr[F[f] + 1]; ++e) {
                const i32 t = g.adj[e];
                if (uidx[t] >= 0) {
                    sidx.push_back(uidx[t]);
                    sw.push_back(g.w[e]);
                    bf.push_back(static_cast<int>(f));
                } else if (fidx[t] >= 0) {
                    S[f * nf + fidx[t]] -= g.w[e];
                }
            }
            sptr[f + 1] = static_cast<long long>(sidx.size());
        }
        uidx.clear();
        uidx.shrink_to_fit();
        fidx.clear();
        fidx.shrink_to_fit();
        DevBuf<double> Z(na > 0 ? static_cast<std::size_t>(na) * nf : 1);
        if (na > 0) {
            SystemSolver sys(solver_precision(cx.precision, static_cast<long long>(nf)), cx.s);
            sys.factor(A);
            A = HostMatrix{};
            const long long tot = na * static_cast<long long>(nf);
            hipLaunchKernelGGL(k_fill_d, dim3(nblk(tot, 256)), dim3(256), 0, cx.s, Z.p, tot, 0.0);
            if (!sidx.empty()) {
                DevBuf<int> dbf = to_device(bf, cx.s), dbe = to_device(sidx, cx.s);
                DevBuf<double> dbw = to_device(sw, cx.s);
                hipLaunchKernelGGL(k_schur_b, dim3(nblk(static_cast<long long>(sidx.size()), 256)), dim3(256), 0,
                                   cx.s, static_cast<long long>(sidx.size()), na, dbf.p, dbe.p, dbw.p, Z.p);
                // Z = A_UU^-1 B, a few right-hand sides at a time, in place
                const int chunk = 16;
                for (std::size_t f0 = 0; f0 < nf; f0 += chunk) {
                    const int kc = static_cast<int>(std::min<std::size_t>(chunk, nf - f0));
                    double* zc = Z.p + static_cast<long long>(f0) * na;
                    sys.solve_device(kc, zc, zc, cx.tol, cx.maxit, cx.st);
                }
                tally(cx, sys);
                DevBuf<long long> dsp = to_device(sptr, cx.s);
                DevBuf<double> dS(nf * nf);
                hipLaunchKernelGGL(k_schur, dim3(nblk(static_cast<long long>(nf) * nf, 256)), dim3(256), 0, cx.s,
                                   static_cast<int>(nf), na, dsp.p, dbe.p, dbw.p, Z.p, dS.p);
                std::vector<double> corr(nf * nf);
                LSC_HIP(hipMemcpyAsync(corr.data(), dS.p, nf * nf * sizeof(double), hipMemcpyDeviceToHost, cx.s));
                LSC_HIP(hipStreamSynchronize(cx.s));
                for (std::size_t t = 0; t < nf * nf; ++t) S[t] += corr[t];  // - A_FU Z, A_FU = -w
            }
        }
        for (std::size_t f = 0; f < nf; ++f) S[f * nf + f] += kShift;
        DevBuf<int> dnodes = to_device(nodes.empty() ? std::vector<i32>{0} : nodes, cx.s);
        DevBuf<int> dF = to_device(F, cx.s);
        DevBuf<double> dyf(nf);
        for (std::size_t p : pk) {
            const Resolved& r = res[p];
            std::vector<std::size_t> U;
            for (std::size_t f = 0; f < nf; ++f)
                if (r.at(r.gnd, F[f]) != kInf) U.push_back(f);
            const std::size_t nu = U.size();
            std::vector<double> Am(nu * nu), y(nu);
            for (std::size_t a = 0; a < nu; ++a) {
                for (std::size_t b = 0; b < nu; ++b) Am[a * nu + b] = S[U[a] * nf + U[b]];
                if (r.any_finite) Am[a * nu + a] += r.at(r.finite, F[U[a]]);
                y[a] = r.at(r.src, F[U[a]]);
            }
            dense_spd_solve(Am, y, nu);
            progress("problem %zu of component %d", p, k);
            std::vector<double> yf(nf, 0.0);
            for (std::size_t a = 0; a < nu; ++a) yf[U[a]] = y[a];
            dyf.upload(yf.data(), nf, cx.s);
            hipLaunchKernelGGL(k_fill_d, dim3(nblk(n, 256)), dim3(256), 0, cx.s, dv.p, static_cast<long long>(n), 0.0);
            if (na > 0)
                hipLaunchKernelGGL(k_schur_v, dim3(nblk(na, 256)), dim3(256), 0, cx.s, na, static_cast<int>(nf), Z.p,
                                   dyf.p, dnodes.p, dv.p);
            hipLaunchKernelGGL(k_scatter_d, dim3(nblk(static_cast<long long>(nf), 256)), dim3(256), 0, cx.s,
                               static_cast<int>(nf), dF.p, dyf.p, dv.p);
            finish(p, k);
        }
    }
    return solved;
}

// Per-problem node voltages and currents on the host (the binding's outputs).
struct NodeSink : Sink {
    const RGraph& g;
    const Members& M;
    hipStream_t s;
    std::vector<NodeSol> out;
    std::vector<double> vh, ch;
    NodeSink(const RGraph& g_, const Members& M_, std::size_t np, hipStream_t s_)
        : g(g_), M(M_), s(s_), out(np), vh(g_.n), ch(g_.n) {
        for (auto& o : out) o.voltage.assign(g.n, 0.0), o.current.assign(g.n, 0.0);
    }
    void component(std::size_t p, i32 k, const Resolved&, const double* d_v, const double* d_cur) override {
        const i32* mem = M.mem.data() + M.start[k];
        const i64 count = M.start[k + 1] - M.start[k];
        LSC_HIP(hipMemcpyAsync(vh.data(), d_v, g.n * sizeof(double), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipMemcpyAsync(ch.data(), d_cur, g.n * sizeof(double), hipMemcpyDeviceToHost, s));
        LSC_HIP(hipStreamSynchronize(s));
        for (i64 q = 0; q < count; ++q) {
            out[p].voltage[mem[q]] = vh[mem[q]];
            out[p].current[mem[q]] = ch[mem[q]];
        }
    }
};

std::vector<NodeSol> advanced_nodes(const RGraph& g, const DevGraph& G, const std::vector<Problem>& probs,
                                    const std::vector<i32>& only, int policy, bool allow_schur, Ctx& cx) {
    const Members M = members_of(g);
    NodeSink sink(g, M, probs.size(), cx.s);
    const std::vector<char> solved = advanced_many(g, G, probs, only, policy, allow_schur, cx, sink);
    for (std::size_t p = 0; p < probs.size(); ++p) {
        sink.out[p].solved = solved[p] != 0;
        const Resolved r = resolve(probs[p], policy);
        for (std::size_t i = 0; i < r.node.size(); ++i)
            if (r.src[i] != 0 || r.gnd[i] != 0) sink.out[p].active.push_back(r.node[i]);
    }
    return std::move(sink.out);
}

// Problems given as cell lists: per-node sums in column-major cell order
// (python.cpp:437-450); the check cell restricts a problem to its component
// (-1: every component; a check cell without a node: none).
void problems_from_cells(const RGraph& g, i64 nprob, const std::int64_t* sptr, const std::int64_t* scells,
                         const double* svals, const std::int64_t* gptr, const std::int64_t* gcells,
                         const double* gvals, const std::int64_t* check, std::vector<Problem>& probs,
                         std::vector<i32>& only) {
    const i64 ncell = g.nx * g.ny;
    probs.assign(nprob, Problem{});
    only.assign(nprob, -1);
    for (i64 p = 0; p < nprob; ++p) {
        // (node, column-major position, value, is-ground)
        std::vector<std::tuple<i32, i64, int, double>> e;
        auto add = [&](const std::int64_t* cells, const double* vals, i64 a, i64 b, int kind) {
            for (i64 k = a; k < b; ++k) {
                const std::int64_t c = cells[k];
                if (c < 0 || c >= ncell) throw std::invalid_argument("source or ground cell out of range");
                const i32 v = g.nodemap[c];
                if (v >= 0 && vals[k] != 0) e.emplace_back(v, (c % g.nx) * g.ny + c / g.nx, kind, vals[k]);
            }
        };
        add(scells, svals, sptr[p], sptr[p + 1], 0);
        add(gcells, gvals, gptr[p], gptr[p + 1], 1);
        std::stable_sort(e.begin(), e.end(), [](const auto& x, const auto& y) {
            return std::make_pair(std::get<0>(x), std::get<1>(x)) < std::make_pair(std::get<0>(y), std::get<1>(y));
        });
        Problem& pb = probs[p];
        for (const auto& [v, pos, kind, val] : e) {
            (void)pos;
            if (pb.node.empty() || pb.node.back() != v) {
                pb.node.push_back(v);
                pb.src.push_back(0.0);
                pb.gnd.push_back(0.0);
            }
            (kind == 0 ? pb.src : pb.gnd).back() += val;
        }
        if (check[p] >= 0) {
            if (check[p] >= ncell) throw std::invalid_argument("check cell out of range");
            const i32 o = g.nodemap[check[p]];
            only[p] = o >= 0 ? o : -2;
        }
    }
}

// cum[c] += cur[node of c], mx[c] = max(mx[c], cur[node of c]) over the cells
// of one component.
__global__ void k_cell_acc(long long cnt, const int* __restrict__ cells, const int* __restrict__ nodemap,
                           const double* __restrict__ cur, double* __restrict__ cum, double* __restrict__ mx) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const int c = cells[i];
    const double x = cur[nodemap[c]];
    cum[c] += x;
    if (mx) mx[c] = fmax(mx[c], x);
}

// Right-hand sides of pairwise jobs, k x n row-major: -1 at the job's source
// unknown, +1 at its target (-1 index: the grounded node, no entry).
__global__ void k_pair_rhs(long long n, int k, const int* __restrict__ a, const int* __restrict__ b,
                           double* __restrict__ B) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n * k) return;
    const int j = static_cast<int>(i / n);
    const int u = static_cast<int>(i - static_cast<long long>(j) * n);
    B[i] = u == a[j] ? -1.0 : (u == b[j] ? 1.0 : 0.0);
}

// out[j * m + i] = X[j * n + u[i]] (0 where u[i] < 0): the solutions at a
// few unknowns, so only those cross PCIe.
__global__ void k_pick(long long n, int k, int m, const int* __restrict__ u, const double* __restrict__ X,
                       double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k * m) return;
    const int j = i / m, q = u[i - j * m];
    out[i] = q >= 0 ? X[static_cast<long long>(j) * n + q] : 0.0;
}

__global__ void k_zero_at(int cnt, const int* __restrict__ idx, double* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) v[idx[i]] = 0.0;
}

__global__ void k_max0(long long n, double* __restrict__ mx) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) mx[i] = fmax(mx[i], 0.0);
}

// The cumulative and maximum current maps over the problems, summed on the
// device per cell in problem order (as run_onetoall sums its per-point maps,
// circuitscape.py:826-848), and each problem's voltage at its check node.
struct SumSink : Sink {
    const RGraph& g;
    hipStream_t s;
    bool focal_zero;
    std::vector<i64> cstart;     // cells of component k: cells[cstart[k] .. cstart[k + 1])
    DevBuf<int> dcells, dnodemap, dz;
    DevBuf<double> cum, mx;
    bool with_max;
    std::vector<i32> check_node;
    std::vector<double> vcheck;
    SumSink(const RGraph& g_, hipStream_t s_, bool focal_zero_, const double* cum_in, const double* max_in,
            std::vector<i32> check_node_)
        : g(g_), s(s_), focal_zero(focal_zero_), with_max(max_in != nullptr), check_node(std::move(check_node_)),
          vcheck(check_node.size(), 0.0) {
        const i64 ncell = g.nx * g.ny;
        cstart.assign(g.ncomp + 1, 0);
        for (i64 c = 0; c < ncell; ++c)
            if (g.nodemap[c] >= 0) ++cstart[g.comp[g.nodemap[c]] + 1];
        for (i32 k = 0; k < g.ncomp; ++k) cstart[k + 1] += cstart[k];
        std::vector<int> cells(cstart[g.ncomp] > 0 ? cstart[g.ncomp] : 1);
        std::vector<i64> fill(cstart.begin(), cstart.end() - 1);
        for (i64 c = 0; c < ncell; ++c)
            if (g.nodemap[c] >= 0) cells[fill[g.comp[g.nodemap[c]]]++] = static_cast<int>(c);
        dcells = to_device(cells, s);
        dnodemap = to_device(g.nodemap, s);
        cum.alloc(ncell ? ncell : 1);
        if (ncell) cum.upload(cum_in, ncell, s);
        if (with_max) {
            mx.alloc(ncell ? ncell : 1);
            if (ncell) mx.upload(max_in, ncell, s);
        }
    }
    void component(std::size_t p, i32 k, const Resolved& r, const double* d_v, const double* d_cur) override {
        double* cur = const_cast<double*>(d_cur);
        if (focal_zero) {
            std::vector<int> act;
            for (std::size_t i = 0; i < r.node.size(); ++i)
                if ((r.src[i] != 0 || r.gnd[i] != 0) && g.comp[r.node[i]] == k) act.push_back(r.node[i]);
            if (!act.empty()) {
                dz = to_device(act, s);
                hipLaunchKernelGGL(k_zero_at, dim3(nblk(static_cast<long long>(act.size()), 256)), dim3(256), 0, s,
                                   static_cast<int>(act.size()), dz.p, cur);
            }
        }
        const long long cnt = cstart[k + 1] - cstart[k];
        if (cnt > 0)
            hipLaunchKernelGGL(k_cell_acc, dim3(nblk(cnt, 256)), dim3(256), 0, s, cnt, dcells.p + cstart[k], dnodemap.p,
                               d_cur, cum.p, with_max ? mx.p : nullptr);
        const i32 v = check_node[p];
        if (v >= 0 && g.comp[v] == k)
            LSC_HIP(hipMemcpyAsync(&vcheck[p], d_v + v, sizeof(double), hipMemcpyDeviceToHost, s));
        if (focal_zero) LSC_HIP(hipStreamSynchronize(s));  // dz is reused
    }
};

std::vector<double> node_sums(const RGraph& g, const double* cells) {
    std::vector<double> s(g.n, 0.0);
    for (i64 x = 0; x < g.nx; ++x)
        for (i64 y = 0; y < g.ny; ++y) {
            const i64 c = y * g.nx + x;
            const i32 v = g.nodemap[c];
            if (v >= 0 && cells[c] != 0) s[v] += cells[c];
        }
    return s;
}

void fill_stats(const Ctx& cx, lsc_solve_stats* out) {
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->ms_fill = cx.st.ms_fill;
    out->ms_factor = cx.st.ms_factor;
    out->ms_solve = cx.st.ms_solve;
    out->n_solved = cx.st.solved;
    out->n_unconverged = cx.st.unconverged;
    out->n_over_1e10 = cx.st.over_bar;
    out->refine_max = cx.st.iter_max;
    out->n_solve_launches = cx.st.solve_launches;
    out->refine_mean = cx.st.solved ? static_cast<double>(cx.st.iter_sum) / cx.st.solved : 0.0;
    out->resid_max = cx.st.resid_max;
    out->resid_mean = cx.st.solved ? cx.st.resid_sum / cx.st.solved : 0.0;
    out->resid_median = cx.st.resid_median();
    out->n_fallback = cx.st.fallback;
}

// A 64-bit digest of arrays, eight bytes at a time: each word mixed by a
// multiply and a shift, then folded in with FNV-1a's offset basis and prime;
// every array's length goes in before its bytes.
struct Digest {
    std::uint64_t h = 14695981039346656037ull;
    void word(std::uint64_t w) {
        w *= 0x9E3779B97F4A7C15ull;
        w ^= w >> 29;
        h = (h ^ w) * 1099511628211ull;
    }
    void bytes(const void* p, std::size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        std::size_t i = 0;
        for (std::uint64_t w; i + 8 <= n; i += 8) {
            std::memcpy(&w, c + i, 8);
            word(w);
        }
        std::uint64_t w = 0;
        std::memcpy(&w, c + i, n - i);
        word(w);
    }
    template <class V>
    void vec(const V& v) {
        word(v.size());
        bytes(v.data(), v.size() * sizeof(v[0]));
    }
};

// The graph's figures and digest, then every component's matrix (no node
// dropped) and plan, timed (lsc_circuitscape_analyze).
void analyze(const RGraph& g, double graph_ms, lsc_analysis* out) {
    lsc_analysis a{};
    a.n_nodes = g.n;
    a.n_components = g.ncomp;
    a.n_entries = g.ptr.empty() ? 0 : g.ptr.back();
    a.graph_ms = graph_ms;
    Digest gd, pd;
    gd.vec(g.nodemap), gd.vec(g.ptr), gd.vec(g.adj), gd.vec(g.w), gd.vec(g.comp);
    auto t0 = std::chrono::steady_clock::now();
    const Members M = members_of(g);
    progress("analysis: %d components' members", g.ncomp);
    std::vector<i32> nodes;
    for (i32 k = 0; k < g.ncomp; ++k) {
        const HostMatrix A = component_matrix(g, M.mem.data() + M.start[k], M.start[k + 1] - M.start[k], nullptr,
                                              [](i32) { return false; }, nodes);
        progress("analysis: component %d matrix of %d unknowns", k, A.n);
        const Plan p = graph_plan(A.bx, A.by, A.cell, A.ptr, A.adj.data(), kGraphLeaf);
        a.plan_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        a.n_fronts += p.nfront;
        a.factor_entries += p.lsize;
        a.largest_front = std::max<int64_t>(a.largest_front, p.max_front);
        a.factor_flops += p.flops;
        pd.vec(p.perm), pd.vec(p.first), pd.vec(p.npiv), pd.vec(p.kids), pd.vec(p.order), pd.vec(p.rows);
        pd.vec(p.relmap), pd.vec(p.fe_ptr), pd.vec(p.fe_row), pd.vec(p.fe_col), pd.vec(p.fe_src), pd.vec(p.lpos);
        t0 = std::chrono::steady_clock::now();
    }
    a.graph_digest = gd.h;
    a.plan_digest = pd.h;
    *out = a;
}

// ---- pairwise (reference raster_pairwise, circuitscape.cpp:474-743)

struct PairOpts {
    bool volt_maps, cur_maps, cum, max, focal_zero, null_currents, null_voltages, log_transform;
    bool network = false;  // a network (no grid): no shortcut; per-pair node and branch currents
    bool maps() const { return volt_maps || cur_maps || cum || max; }
};

// Branch currents of one component (reference network_currents,
// circuitscape.cpp:340-362): every edge once, from its larger node c to its
// smaller neighbour r, |w| (v_r - v_c); values below 1e-8 of the largest
// set to 0, then all made positive. Per branch: its adjacency entry, its
// nodes r < c and its current.
struct Branches {
    std::vector<i64> e;
    std::vector<i32> r, c;
    std::vector<double> cur;
};
void branch_currents(const RGraph& g, const i32* mem, i64 count, const double* v, Branches& b) {
    b.e.clear(), b.r.clear(), b.c.clear(), b.cur.clear();
    double top = -kInf;
    for (i64 p = 0; p < count; ++p) {
        const i32 c = mem[p];
        for (i64 e = g.ptr[c]; e < g.ptr[c + 1]; ++e) {
            const i32 r = g.adj[e];
            if (r >= c) continue;
            const double x = std::fabs(g.w[e]) * (v[r] - v[c]);
            top = std::max(top, x);
            b.e.push_back(e), b.r.push_back(r), b.c.push_back(c), b.cur.push_back(x);
        }
    }
    for (double& x : b.cur) {
        if (std::fabs(x / top) < 1e-8) x = 0.0;
        x = std::fabs(x);
    }
}

// What a network pairwise run returns besides the resistances: per-pair
// tables through `emit`, and with any map option the node currents summed
// by node and the branch currents summed by adjacency entry.
// Job j's voltages relative to its source, by graph node, over one
// component's members (unknown u of node c: uidx[c]; the grounded node 0).
__global__ void k_job_voltages(int cnt, const int* __restrict__ mem, const int* __restrict__ uidx,
                               const double* __restrict__ Xj, double vs, double* __restrict__ vnode) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cnt) return;
    const int c = mem[p], u = uidx[c];
    vnode[c] = (u >= 0 ? Xj[u] : 0.0) - vs;
}

// One pair's maps on the cells (the raster pairwise's postprocessing): the
// voltage map (vmap, optional), the current map (cmap, optional) with the
// focal-zero, log10 and null options, added `times` times into cum and
// maximised into mx (each when given), one thread per cell, so the sums run
// in job order as on a single thread. flags: 1 focal currents zero, 2 log10,
// 4 null currents, 8 null voltages (nodata where the conductance is 0).
__global__ void k_pair_cells(long long ncell, const int* __restrict__ nodemap, const int* __restrict__ comp, int k,
                             const double* __restrict__ vnode, const double* __restrict__ cnode,
                             const unsigned char* __restrict__ zero_g, int flags, int na, int nb, int times,
                             double* __restrict__ cum, double* __restrict__ mx, double* __restrict__ vmap,
                             double* __restrict__ cmap) {
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int v = nodemap[c];
    const bool in = v >= 0 && comp[v] == k;
    const bool null_cell = zero_g && zero_g[c];
    if (vmap) vmap[c] = (flags & 8) && null_cell ? -9999.0 : (in ? vnode[v] : 0.0);
    double x = in ? cnode[v] : 0.0;
    if (in && (flags & 1) && (v == na || v == nb)) x = 0.0;
    if (flags & 2) x = x > 0 ? log10(x) : -9999.0;
    if ((flags & 4) && null_cell) x = -9999.0;
    if (cmap) cmap[c] = x;
    if (cum)
        for (int t = 0; t < times; ++t) cum[c] += x;
    if (mx && times > 0) mx[c] = mx[c] < x ? x : mx[c];
}

__global__ void k_fill_value(long long n, double value, double* __restrict__ a) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) a[q] = value;
}

// Right-hand sides solved at once in pairwise mode: as many as a dozen
// vectors of the component's size each fit in half the free device memory,
// at most 1024 (LSC_PAIR_CHUNK lowers the cap; the results do not depend on
// it: every right-hand side is solved on its own and the maps add in job order).
long long pair_chunk(long long jobs, long long n) {
    std::size_t free_b = 0, total_b = 0;
    LSC_HIP(hipMemGetInfo(&free_b, &total_b));
    const double per_rhs = 16.0 * 8.0 * static_cast<double>(std::max<long long>(n, 1));
    long long cap = 1024;
    if (const char* e = std::getenv("LSC_PAIR_CHUNK"))
        if (std::atoi(e) > 0) cap = std::min<long long>(cap, std::atoi(e));
    const long long fit = static_cast<long long>(0.5 * static_cast<double>(free_b) / per_rhs);
    return std::max<long long>(1, std::min<long long>({jobs, fit, cap}));
}

struct NetPairs {
    lsc_network_emit_fn emit = nullptr;
    void* ctx = nullptr;
    std::vector<double> cum, cum_branch;
};

void pairwise(const RGraph& g, const DevGraph& G, const double* gmap, const std::vector<i32>& points,
              const std::uint8_t* excluded, const PairOpts& o, lsc_emit_fn emit, void* ctx, double* R_out,
              double* cum_out, double* max_out, Ctx& cx, NetPairs* net = nullptr) {
    const i32 np = static_cast<i32>(points.size());
    const i64 ncell = g.nx * g.ny;
    std::vector<double> Rm(static_cast<std::size_t>(np) * np, -1.0);
    auto R = [&](i32 a, i32 b) -> double& { return Rm[static_cast<std::size_t>(a) * np + b]; };
    bool any_ex = false;
    for (i64 k = 0; k < static_cast<i64>(np) * np; ++k) any_ex = any_ex || excluded[k];
    const bool shortcut = !o.network && !o.maps() && !any_ex;
    std::vector<double> voltmatrix, sres;
    if (shortcut) voltmatrix.assign(static_cast<std::size_t>(np) * np, 0.0), sres.assign(static_cast<std::size_t>(np) * np, -1.0);
    if (o.network && o.maps()) net->cum.assign(g.n, 0.0), net->cum_branch.assign(g.adj.size(), 0.0);
    const Members M = members_of(g);
    Branches branch;
    std::vector<double> cur_mem;
    std::unordered_map<i32, std::vector<i32>> positions;
    for (i32 i = 0; i < np; ++i)
        if (points[i] >= 0) positions[points[i]].push_back(i);
    DevBuf<int> dmem = to_device(M.mem.empty() ? std::vector<i32>{0} : M.mem, cx.s);
    DevBuf<double> dv, dcur, scratch;
    // raster maps on the device: node of every cell, component of every node,
    // cells of conductance 0 (for the null options), the running cum / max
    // maps, one pair's maps when they are emitted
    DevBuf<int> dnodemap, dcomp;
    DevBuf<unsigned char> dzero;
    DevBuf<double> dcum, dmx, dvmap, dcmap;
    const int cell_flags = (o.focal_zero ? 1 : 0) | (o.log_transform ? 2 : 0) | (o.null_currents ? 4 : 0) |
                           (o.null_voltages ? 8 : 0);
    if (o.maps() && !o.network) {
        dnodemap = to_device(g.nodemap.empty() ? std::vector<i32>{-1} : g.nodemap, cx.s);
        dcomp = to_device(g.comp.empty() ? std::vector<i32>{0} : g.comp, cx.s);
        if (o.null_currents || o.null_voltages) {
            std::vector<unsigned char> z(static_cast<std::size_t>(ncell));
            for (i64 c = 0; c < ncell; ++c) z[c] = gmap[c] == 0;
            dzero = to_device(z.empty() ? std::vector<unsigned char>{0} : z, cx.s);
        }
        const unsigned cb = nblk(ncell, 256);
        if (o.cum) {
            dcum.alloc(ncell ? ncell : 1);
            if (ncell) hipLaunchKernelGGL(k_fill_value, dim3(cb), dim3(256), 0, cx.s, ncell, 0.0, dcum.p);
        }
        if (o.max) {
            dmx.alloc(ncell ? ncell : 1);
            if (ncell) hipLaunchKernelGGL(k_fill_value, dim3(cb), dim3(256), 0, cx.s, ncell, -9999.0, dmx.p);
        }
    }
    std::vector<i32> nodes;
    for (i32 k = 0; k < g.ncomp; ++k) {
        const i64 p0 = M.start[k], p1 = M.start[k + 1];
        std::vector<i32> csub;
        {
            std::unordered_set<i32> seen;
            for (i32 i = 0; i < np; ++i) {
                const i32 v = points[i];
                if (v >= 0 && g.comp[v] == k && seen.insert(v).second) csub.push_back(v);
            }
        }
        if (csub.empty()) continue;
        for (i32 v : csub) {
            const auto& pos = positions[v];
            for (std::size_t a = 0; a < pos.size(); ++a)
                for (std::size_t b = a + 1; b < pos.size(); ++b) R(pos[a], pos[b]) = R(pos[b], pos[a]) = 0.0;
        }
        std::vector<std::pair<i32, i32>> jobs;
        const std::size_t nsrc = shortcut ? std::min<std::size_t>(1, csub.size()) : csub.size();
        for (std::size_t si = 0; si < nsrc; ++si)
            for (std::size_t di = si + 1; di < csub.size(); ++di) {
                bool any = false;
                for (i32 ci : positions[csub[si]])
                    for (i32 cj : positions[csub[di]]) any = any || !excluded[static_cast<std::size_t>(ci) * np + cj];
                if (any) jobs.push_back({csub[si], csub[di]});
            }
        if (jobs.empty()) continue;
        // one node grounded (the component's last): exact voltage differences
        const i32 ref = M.mem[p1 - 1];
        const i64 count = p1 - p0;
        const i32* mb = M.mem.data() + p0;
        progress("pairwise: component %d matrix (%lld nodes)", k, static_cast<long long>(count));
        HostMatrix A = component_matrix(g, mb, count, nullptr, [&](i32 v) { return v == ref; }, nodes);
        SystemSolver S(solver_precision(cx.precision, static_cast<long long>(jobs.size())), cx.s);
        S.factor(A);
        progress("pairwise: factored; %zu right-hand sides", jobs.size());
        std::vector<i32> uidx(g.n, -1);  // node -> unknown (-1: the grounded node, other components)
        for (int i = 0; i < A.n; ++i) uidx[nodes[i]] = i;
        const std::size_t nj = jobs.size();
        const int nfc = static_cast<int>(csub.size());
        std::unordered_map<i32, int> fpos;  // focal node -> column of the picked solutions
        std::vector<int> fu(nfc);
        for (int i = 0; i < nfc; ++i) fu[i] = uidx[csub[i]], fpos[csub[i]] = i;
        const bool raster_maps = o.maps() && !o.network;
        // the raster maps stay on the device: the component's unknowns by node,
        // the job's node voltages and currents
        DevBuf<int> duidx, dfu = to_device(fu, cx.s);
        if (raster_maps) {
            duidx = to_device(uidx, cx.s);
            dv.alloc(g.n ? g.n : 1), dcur.alloc(g.n ? g.n : 1);
        }
        std::vector<double> vnode, cnode, vcol;
        if (o.network && o.maps()) vnode.assign(g.n, 0.0), cnode.assign(g.n, 0.0), vcol.resize(count);
        if (o.network && o.maps()) dv.alloc(g.n ? g.n : 1), dcur.alloc(g.n ? g.n : 1);
        // jobs in chunks whose right-hand sides fit the device (all at once when they do)
        const long long chunk = A.n > 0 ? pair_chunk(static_cast<long long>(nj), A.n) : static_cast<long long>(nj);
        DevBuf<double> dB, dX, dXf;
        std::vector<double> Xf, X;
        for (std::size_t j0 = 0; j0 < nj; j0 += static_cast<std::size_t>(chunk)) {
            const int kc = static_cast<int>(std::min<std::size_t>(static_cast<std::size_t>(chunk), nj - j0));
            Xf.assign(static_cast<std::size_t>(kc) * nfc, 0.0);
            if (A.n > 0) {
                std::vector<int> ja(kc), jb(kc);
                for (int j = 0; j < kc; ++j) ja[j] = uidx[jobs[j0 + j].first], jb[j] = uidx[jobs[j0 + j].second];
                DevBuf<int> dja = to_device(ja, cx.s), djb = to_device(jb, cx.s);
                const long long tot = static_cast<long long>(kc) * A.n;
                if (dB.n < static_cast<std::size_t>(tot)) dB.alloc(tot), dX.alloc(tot);
                if (dXf.n < Xf.size()) dXf.alloc(Xf.size());
                hipLaunchKernelGGL(k_pair_rhs, dim3(static_cast<unsigned>((tot + 255) / 256)), dim3(256), 0, cx.s,
                                   static_cast<long long>(A.n), kc, dja.p, djb.p, dB.p);
                S.solve_device(kc, dB.p, dX.p, cx.tol, cx.maxit, cx.st);
                hipLaunchKernelGGL(k_pick, dim3(static_cast<unsigned>((Xf.size() + 255) / 256)), dim3(256), 0, cx.s,
                                   static_cast<long long>(A.n), kc, nfc, dfu.p, dX.p, dXf.p);
                LSC_HIP(hipMemcpyAsync(Xf.data(), dXf.p, Xf.size() * sizeof(double), hipMemcpyDeviceToHost, cx.s));
                if (o.network && o.maps()) {
                    X.resize(static_cast<std::size_t>(tot));
                    LSC_HIP(hipMemcpyAsync(X.data(), dX.p, tot * sizeof(double), hipMemcpyDeviceToHost, cx.s));
                }
                LSC_HIP(hipStreamSynchronize(cx.s));
            } else if (o.network && o.maps()) {
                X.assign(static_cast<std::size_t>(kc) * A.n, 0.0);
            }
            // voltage of focal node v in job j of the chunk (relative to the grounded node)
            auto fvolt = [&](int j, i32 v) { return Xf[static_cast<std::size_t>(j) * nfc + fpos.at(v)]; };
            for (int j = 0; j < kc; ++j) {
                const auto& job = jobs[j0 + j];
                const double vs = fvolt(j, job.first);
                const double resistance = (fvolt(j, job.second) - vs) - (vs - vs);
                int times = 0;  // the job's position pairs that are not excluded
                for (i32 ci : positions[job.first])
                    for (i32 cj : positions[job.second]) {
                        if (excluded[static_cast<std::size_t>(ci) * np + cj]) continue;
                        ++times;
                        R(ci, cj) = R(cj, ci) = resistance;
                        if (!shortcut) continue;
                        for (i32 i = 1; i < np; ++i) {
                            const i32 v = points[i];
                            if (v < 0 || g.comp[v] != k) continue;
                            voltmatrix[static_cast<std::size_t>(i) * np + cj] = 1 - ((fvolt(j, v) - vs) / resistance);
                        }
                    }
                if (!o.maps() || times == 0) continue;
                if (raster_maps) {
                    const double* Xj = A.n > 0 ? dX.p + static_cast<std::size_t>(j) * A.n : nullptr;
                    if (A.n > 0)
                        hipLaunchKernelGGL(k_job_voltages, dim3(nblk(count, 256)), dim3(256), 0, cx.s,
                                           static_cast<int>(count), dmem.p + p0, duidx.p, Xj, vs, dv.p);
                    else
                        hipLaunchKernelGGL(k_fill_value, dim3(nblk(g.n, 256)), dim3(256), 0, cx.s,
                                           static_cast<long long>(g.n), 0.0, dv.p);
                    graph_currents(G, dmem.p + p0, static_cast<int>(count), dv.p, nullptr, dcur.p, scratch, cx.s);
                    const bool want_emit = emit && (o.volt_maps || o.cur_maps);
                    if (want_emit) {
                        if (o.volt_maps && dvmap.n < static_cast<std::size_t>(ncell)) dvmap.alloc(ncell);
                        if (dcmap.n < static_cast<std::size_t>(ncell)) dcmap.alloc(ncell);
                    }
                    hipLaunchKernelGGL(k_pair_cells, dim3(nblk(ncell, 256)), dim3(256), 0, cx.s, ncell, dnodemap.p,
                                       dcomp.p, k, dv.p, dcur.p, dzero.p, cell_flags, job.first, job.second, times,
                                       o.cum ? dcum.p : nullptr, o.max ? dmx.p : nullptr,
                                       want_emit && o.volt_maps ? dvmap.p : nullptr, want_emit ? dcmap.p : nullptr);
                    if (!want_emit) continue;
                    std::vector<double> vmap(o.volt_maps ? ncell : 0), cmap(ncell);
                    if (o.volt_maps)
                        LSC_HIP(hipMemcpyAsync(vmap.data(), dvmap.p, ncell * sizeof(double), hipMemcpyDeviceToHost, cx.s));
                    LSC_HIP(hipMemcpyAsync(cmap.data(), dcmap.p, ncell * sizeof(double), hipMemcpyDeviceToHost, cx.s));
                    LSC_HIP(hipStreamSynchronize(cx.s));
                    for (i32 ci : positions[job.first])
                        for (i32 cj : positions[job.second])
                            if (!excluded[static_cast<std::size_t>(ci) * np + cj])
                                emit(ci, cj, o.volt_maps ? vmap.data() : nullptr, o.cur_maps ? cmap.data() : nullptr, ctx);
                    continue;
                }
                // network: the component's node voltages, node and branch currents per pair, on the host
                for (i64 q = 0; q < count; ++q) {
                    const i32 u = uidx[mb[q]];
                    vcol[q] = (u >= 0 ? X[static_cast<std::size_t>(j) * A.n + u] : 0.0) - vs;
                    vnode[mb[q]] = vcol[q];
                }
                dv.upload(vnode.data(), g.n, cx.s);
                graph_currents(G, dmem.p + p0, static_cast<int>(count), dv.p, nullptr, dcur.p, scratch, cx.s);
                LSC_HIP(hipMemcpyAsync(cnode.data(), dcur.p, g.n * sizeof(double), hipMemcpyDeviceToHost, cx.s));
                LSC_HIP(hipStreamSynchronize(cx.s));
                branch_currents(g, mb, count, vnode.data(), branch);
                cur_mem.resize(count);
                for (i32 ci : positions[job.first])
                    for (i32 cj : positions[job.second]) {
                        if (excluded[static_cast<std::size_t>(ci) * np + cj]) continue;
                        for (i64 q = 0; q < count; ++q) {
                            cur_mem[q] = cnode[mb[q]];
                            net->cum[mb[q]] += cnode[mb[q]];
                        }
                        for (std::size_t b = 0; b < branch.e.size(); ++b) net->cum_branch[branch.e[b]] += branch.cur[b];
                        if (net->emit)
                            net->emit(ci, cj, count, mb, o.volt_maps ? vcol.data() : nullptr, cur_mem.data(),
                                      static_cast<i64>(branch.e.size()), branch.r.data(), branch.c.data(),
                                      branch.cur.data(), net->ctx);
                    }
            }
        }
        progress("pairwise: solved");
        tally(cx, S);
        if (shortcut) {
            // update_shortcut_resistances! (circuitscape.cpp:708-734)
            const i32 anchor = positions[csub[0]].front();
            auto Sr = [&](i32 a, i32 b) -> double& { return sres[static_cast<std::size_t>(a) * np + b]; };
            constexpr double kInvalid = -777;
            std::vector<std::uint8_t> check(np, 0);
            for (i32 i = 0; i < np; ++i) check[i] = points[i] >= 0 && g.comp[points[i]] == k;
            for (i32 x = 0; x < np; ++x) {
                if (!check[x]) continue;
                const double r1x = R(anchor, x);
                if (r1x == -1) continue;
                Sr(x, anchor) = Sr(anchor, x) = r1x;
                for (i32 y = x; y < np; ++y) {
                    if (!check[y]) continue;
                    const double r12 = R(anchor, y);
                    if (r12 == -1) continue;
                    if (r1x != kInvalid) {
                        Sr(anchor, y) = Sr(y, anchor) = r12;
                        const double vx = voltmatrix[static_cast<std::size_t>(x) * np + y];
                        const double r2x = 2 * r12 * vx + r1x - r12;
                        if (Sr(y, x) != kInvalid) Sr(y, x) = Sr(x, y) = r2x;
                    } else {
                        for (i32 t = 0; t < np; ++t) Sr(x, t) = Sr(t, x) = kInvalid;
                    }
                }
            }
        }
    }
    if (shortcut) Rm = std::move(sres);
    for (i32 i = 0; i < np; ++i) Rm[static_cast<std::size_t>(i) * np + i] = 0.0;
    std::copy(Rm.begin(), Rm.end(), R_out);
    if (o.network) return;
    // the maps at -9999 or above, as the reference writes them
    auto fetch = [&](const DevBuf<double>& d, double* out) {
        LSC_HIP(hipMemcpyAsync(out, d.p, ncell * sizeof(double), hipMemcpyDeviceToHost, cx.s));
        LSC_HIP(hipStreamSynchronize(cx.s));
        for (i64 c = 0; c < ncell; ++c) out[c] = out[c] < -9999 ? -9999 : out[c];
    };
    if (o.cum && cum_out && ncell) fetch(dcum, cum_out);
    if (o.max && max_out && ncell) fetch(dmx, max_out);
}

// ---- networks (the reference's network_advanced / network_pairwise,
// src/cpp/circuit/python.cpp:310-402)

// Advanced mode on a network: voltages by node (0 where nothing was solved),
// then the node currents of the whole graph with the finite grounds (taken
// before the policy) and its branch currents.
void network_advanced(const RGraph& g, const DevGraph& G, const double* source, const double* ground, int policy,
                      Ctx& cx, double* voltage, double* current, Branches& br, bool& solved) {
    const i32 m = g.n;
    std::vector<double> src(source, source + m), gnd(ground, ground + m), finite(m, 0.0);
    bool any_finite = false;
    for (i32 v = 0; v < m; ++v) {
        finite[v] = gnd[v] < kInf ? gnd[v] : 0.0;
        any_finite = any_finite || finite[v] != 0;
    }
    auto sol = advanced_nodes(g, G, {problem_from_dense(src, gnd)}, {-1}, policy, false, cx);
    std::vector<i32> all(m);
    for (i32 v = 0; v < m; ++v) all[v] = v;
    DevBuf<int> dall = to_device(all, cx.s);
    DevBuf<double> dv = to_device(sol[0].voltage, cx.s), dfin = to_device(finite, cx.s), dcur, scratch;
    dcur.alloc(m);
    graph_currents(G, dall.p, m, dv.p, any_finite ? dfin.p : nullptr, dcur.p, scratch, cx.s);
    LSC_HIP(hipMemcpyAsync(current, dcur.p, m * sizeof(double), hipMemcpyDeviceToHost, cx.s));
    LSC_HIP(hipStreamSynchronize(cx.s));
    std::copy(sol[0].voltage.begin(), sol[0].voltage.end(), voltage);
    branch_currents(g, all.data(), m, sol[0].voltage.data(), br);
    solved = sol[0].solved;
}

// ---- the adjoint of a grid circuit

// dJ/dr per cell: every edge e = (a, b) adds -(sum_k (lam_a - lam_b)(v_a - v_b)) dw_e/dr_c at both ends,
// w_e = (g_a + g_b) c_e, g = 1 / r, so dw_e/dr_c = -c_e / r_c^2.
__global__ void k_adjoint_grad(long long nx, long long ny, int four, int npairs, const double* __restrict__ r,
                               const double* __restrict__ V, const double* __restrict__ Lam,
                               double* __restrict__ grad) {
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long n = nx * ny;
    if (c >= n) return;
    const double rc = r[c];
    if (!(rc > 0) || !isfinite(rc)) {
        grad[c] = 0.0;
        return;
    }
    const long long y = c / nx, x = c - y * nx;
    const int nd = four ? 4 : 8;
    const int dy8[8] = {-1, 0, 1, -1, 1, -1, 0, 1}, dx8[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    const int dy4[4] = {0, -1, 1, 0}, dx4[4] = {-1, 0, 0, 1};
    double acc = 0.0;
    for (int d = 0; d < nd; ++d) {
        const int dy = four ? dy4[d] : dy8[d], dx = four ? dx4[d] : dx8[d];
        const long long yy = y + dy, xx = x + dx;
        if (yy < 0 || yy >= ny || xx < 0 || xx >= nx) continue;
        const long long q = yy * nx + xx;
        const double rq = r[q];
        if (!(rq > 0) || !isfinite(rq)) continue;
        const double ce = (dy != 0 && dx != 0) ? 1.0 / 2.8284271247461903 : 0.5;
        double s = 0.0;
        for (int k = 0; k < npairs; ++k) {
            const long long o = static_cast<long long>(k) * n;
            s += (Lam[o + c] - Lam[o + q]) * (V[o + c] - V[o + q]);
        }
        acc += s * ce;
    }
    grad[c] = acc / (rc * rc);
}

}  // namespace

// Right-hand sides by cell to a component's unknowns (k x nu), and back.
__global__ void k_cells_to_nodes(int nu, int k, long long ncell, const long long* __restrict__ cell,
                                 const double* __restrict__ by_cell, double* __restrict__ by_node) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= static_cast<long long>(nu) * k) return;
    const long long j = q / nu, i = q - j * nu;
    by_node[q] = by_cell[j * ncell + cell[i]];
}

__global__ void k_nodes_to_cells(int nu, int k, long long ncell, const long long* __restrict__ cell,
                                 const double* __restrict__ by_node, double* __restrict__ by_cell) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= static_cast<long long>(nu) * k) return;
    const long long j = q / nu, i = q - j * nu;
    by_cell[j * ncell + cell[i]] = by_node[q];
}

__global__ void k_fill_zero(long long n, double* __restrict__ a) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < n) a[q] = 0.0;
}

// A grid circuit kept on the GPU for gradients: conductance 1 / r on the
// cells with a finite r > 0 (8 or 4 neighbours, average conductance), every
// component grounded at its last node, factorised once; new resistances on
// the same cells are factorised again on the same symbolic plans.
struct GridCircuit {
    Ctx cx;
    RGraph g;
    Members M;
    std::vector<std::unique_ptr<SystemSolver>> sys;  // per component
    std::vector<std::vector<i32>> nodes;             // unknown -> graph node, per component
    std::vector<DevBuf<long long>> dcell;            // unknown -> cell, per component (device)
    std::vector<unsigned char> valid;                // cells in the circuit (finite r > 0)
    DevBuf<double> r;                                // resistance (cells)
    DevBuf<double> B, X;                             // right-hand sides and solutions by unknown
    long long nx = 0, ny = 0;
    int four = 0;
};

bool valid_resistance(double v) { return v > 0 && std::isfinite(v); }

// The component matrices of the grid of conductance 1 / r: each component
// grounded at its last member, factorised (`refactor`: on the plans of an
// earlier factorisation of the same cells).
void grid_factor(GridCircuit& C, const double* resistance, bool refactor) {
    const i64 n = C.nx * C.ny;
    std::vector<double> g(n);
    for (i64 c = 0; c < n; ++c) g[c] = valid_resistance(resistance[c]) ? 1 / resistance[c] : 0.0;
    C.g = raster_graph(g.data(), nullptr, C.nx, C.ny, false, C.four != 0);
    C.M = members_of(C.g);
    if (refactor && static_cast<std::size_t>(C.g.ncomp) != C.sys.size())
        throw std::invalid_argument("the circuit's components changed");
    for (i32 k = 0; k < C.g.ncomp; ++k) {
        const i64 p0 = C.M.start[k], p1 = C.M.start[k + 1];
        const i32 ref = C.M.mem[p1 - 1];
        std::vector<i32> nodes;
        HostMatrix A = component_matrix(C.g, C.M.mem.data() + p0, p1 - p0, nullptr, [&](i32 v) { return v == ref; },
                                        nodes);
        if (refactor) {
            if (nodes != C.nodes[k]) throw std::invalid_argument("the circuit's components changed");
            C.sys[k]->refactor(A);
            continue;
        }
        // a grid circuit is kept for repeated solves (the adjoint's pairs)
        auto S = std::make_unique<SystemSolver>(solver_precision(C.cx.precision, kAutoFp64Rhs), C.cx.s);
        S->factor(A);
        std::vector<long long> cells(nodes.size());
        for (std::size_t i = 0; i < nodes.size(); ++i) cells[i] = C.g.node_cell[nodes[i]];
        C.dcell.push_back(to_device(cells.empty() ? std::vector<long long>{0} : cells, C.cx.s));
        C.sys.push_back(std::move(S));
        C.nodes.push_back(std::move(nodes));
    }
    C.r.alloc(n ? n : 1);
    if (n) C.r.upload(resistance, n, C.cx.s);
    LSC_HIP(hipStreamSynchronize(C.cx.s));
}

}  // namespace lsc

namespace {

using lsc::i32;
using lsc::i64;

// Potentials by cell (device, k x ny x nx) of right-hand sides given by cell
// (device, k x ny x nx): every component solved on the device, cells outside
// the circuit 0.
void grid_solve(lsc::GridCircuit& C, int k, const double* d_rhs_cells, double* d_out_cells) {
    const i64 ncell = C.nx * C.ny;
    const long long tot = static_cast<long long>(k) * ncell;
    if (tot > 0)
        hipLaunchKernelGGL(lsc::k_fill_zero, dim3(lsc::nblk(tot, 256)), dim3(256), 0, C.cx.s, tot, d_out_cells);
    for (std::size_t ci = 0; ci < C.sys.size(); ++ci) {
        auto& S = *C.sys[ci];
        const int nu = S.n();
        if (nu == 0 || k <= 0) continue;
        const std::size_t kn = static_cast<std::size_t>(k) * nu;
        if (C.B.n < kn) C.B.alloc(kn), C.X.alloc(kn);
        const long long* cell = C.dcell[ci].p;
        hipLaunchKernelGGL(lsc::k_cells_to_nodes, dim3(lsc::nblk(static_cast<long long>(kn), 256)), dim3(256), 0,
                           C.cx.s, nu, k, static_cast<long long>(ncell), cell, d_rhs_cells, C.B.p);
        S.solve_device(k, C.B.p, C.X.p, C.cx.tol, C.cx.maxit, C.cx.st);
        hipLaunchKernelGGL(lsc::k_nodes_to_cells, dim3(lsc::nblk(static_cast<long long>(kn), 256)), dim3(256), 0,
                           C.cx.s, nu, k, static_cast<long long>(ncell), cell, C.X.p, d_out_cells);
    }
    LSC_HIP(hipStreamSynchronize(C.cx.s));
}

}  // namespace

extern "C" {

int lsc_circuit_advanced(const double* conductance, const double* source, const double* ground, int64_t nx,
                         int64_t ny, int four_neighbors, const lsc_solve_options* options, double* current,
                         double* voltage, lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        if (nx < 0 || ny < 0) throw std::invalid_argument("negative raster size");
        lsc::Ctx cx = lsc::make_ctx(options);
        const i64 n = nx * ny;
        // cells with a conductance > 0 are nodes (grid.cpp:83-106)
        std::vector<double> g(n);
        for (i64 c = 0; c < n; ++c) g[c] = conductance[c] > 0 ? conductance[c] : 0.0;
        lsc::RGraph G = lsc::raster_graph(g.data(), nullptr, nx, ny, false, four_neighbors != 0);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        std::vector<double> src(G.n, 0.0), gnd(G.n, 0.0);
        for (i64 c = 0; c < n; ++c) {
            const i32 v = G.nodemap[c];
            if (v < 0) continue;
            src[v] = source ? source[c] : 0.0;
            gnd[v] = ground ? ground[c] : 0.0;
        }
        // a source on a ground is dropped (grid.cpp:110-123): the rmvsrc policy
        auto sol = lsc::advanced_nodes(G, DG, {lsc::problem_from_dense(src, gnd)}, {-1}, 1, false, cx);
        for (i64 c = 0; c < n; ++c) {
            const i32 v = G.nodemap[c];
            current[c] = v >= 0 ? sol[0].current[v] : 0.0;
            voltage[c] = v >= 0 ? sol[0].voltage[v] : std::numeric_limits<double>::quiet_NaN();
        }
        lsc::fill_stats(cx, stats);
    });
}

int lsc_circuitscape_advanced(const double* gmap, const int64_t* polymap, const double* source, const double* ground,
                              int64_t nx, int64_t ny, int avg_resistances, int four_neighbors, int policy,
                              const lsc_solve_options* options, int32_t* nodemap, double* voltage, double* current,
                              int64_t* n_nodes, int32_t* solved, int32_t* n_components, lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        if (policy < 0 || policy > 3) throw std::invalid_argument("remove_src_or_gnd must be keepall, rmvsrc, rmvgnd or rmvall");
        lsc::Ctx cx = lsc::make_ctx(options);
        lsc::RGraph G = lsc::raster_graph(gmap, polymap, nx, ny, avg_resistances != 0, four_neighbors != 0);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        auto sol = lsc::advanced_nodes(
            G, DG, {lsc::problem_from_dense(lsc::node_sums(G, source), lsc::node_sums(G, ground))}, {-1}, policy, false,
            cx);
        for (i64 c = 0; c < nx * ny; ++c) nodemap[c] = G.nodemap[c];
        std::copy(sol[0].voltage.begin(), sol[0].voltage.end(), voltage);
        std::copy(sol[0].current.begin(), sol[0].current.end(), current);
        *n_nodes = G.n;
        *solved = sol[0].solved;
        *n_components = G.ncomp;
        lsc::fill_stats(cx, stats);
    });
}

int lsc_circuitscape_pairwise(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                              const int64_t* point_cells, int64_t npts, const uint8_t* excluded, int avg_resistances,
                              int four_neighbors, int flags, lsc_emit_fn emit, void* ctx,
                              const lsc_solve_options* options, double* resistances, double* cum, double* max,
                              lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        lsc::Ctx cx = lsc::make_ctx(options);
        lsc::progress("pairwise: raster graph of %lld x %lld cells", static_cast<long long>(nx), static_cast<long long>(ny));
        lsc::RGraph G = lsc::raster_graph(gmap, polymap, nx, ny, avg_resistances != 0, four_neighbors != 0);
        lsc::progress("pairwise: %d nodes, %d components; upload", G.n, G.ncomp);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        std::vector<i32> nodes(npts);
        for (int64_t i = 0; i < npts; ++i) {
            const int64_t c = point_cells[i];
            nodes[i] = c >= 0 && c < nx * ny ? G.nodemap[c] : -1;
        }
        lsc::PairOpts o{(flags & 1) != 0,  (flags & 2) != 0,  (flags & 4) != 0,  (flags & 8) != 0,
                        (flags & 16) != 0, (flags & 32) != 0, (flags & 64) != 0, (flags & 128) != 0};
        lsc::pairwise(G, DG, gmap, nodes, excluded, o, emit, ctx, resistances, cum, max, cx);
        lsc::fill_stats(cx, stats);
    });
}

int lsc_network_advanced(const int32_t* a, const int32_t* b, const double* w, int64_t n_edges, int32_t n_nodes,
                         const double* source, const double* ground, int policy, const lsc_solve_options* options,
                         double* voltage, double* current, int32_t* branch_r, int32_t* branch_c,
                         double* branch_current, int64_t* n_branches, int32_t* solved, int32_t* n_components,
                         lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        if (policy < 0 || policy > 3) throw std::invalid_argument("remove_src_or_gnd must be keepall, rmvsrc, rmvgnd or rmvall");
        if (n_nodes <= 0 || n_edges < 0) throw std::invalid_argument("a network needs nodes");
        lsc::Ctx cx = lsc::make_ctx(options);
        lsc::RGraph G = lsc::network_graph(a, b, w, n_edges, n_nodes);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        lsc::Branches br;
        bool ok = false;
        lsc::network_advanced(G, DG, source, ground, policy, cx, voltage, current, br, ok);
        if (static_cast<int64_t>(br.e.size()) > n_edges) throw std::logic_error("more branches than edges");
        std::copy(br.r.begin(), br.r.end(), branch_r);
        std::copy(br.c.begin(), br.c.end(), branch_c);
        std::copy(br.cur.begin(), br.cur.end(), branch_current);
        *n_branches = static_cast<int64_t>(br.e.size());
        *solved = ok;
        *n_components = G.ncomp;
        lsc::fill_stats(cx, stats);
    });
}

int lsc_network_pairwise(const int32_t* a, const int32_t* b, const double* w, int64_t n_edges, int32_t n_nodes,
                         const int32_t* points, int64_t n_points, const uint8_t* excluded, int flags,
                         lsc_network_emit_fn emit, void* ctx, const lsc_solve_options* options, double* resistances,
                         double* cum, int32_t* cum_branch_r, int32_t* cum_branch_c, double* cum_branch_current,
                         int64_t* n_cum_branches, int32_t* n_components, lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        if (n_nodes <= 0 || n_edges < 0) throw std::invalid_argument("a network needs nodes");
        lsc::Ctx cx = lsc::make_ctx(options);
        lsc::RGraph G = lsc::network_graph(a, b, w, n_edges, n_nodes);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        std::vector<lsc::i32> pts(n_points);
        for (int64_t i = 0; i < n_points; ++i) pts[i] = points[i] >= 0 && points[i] < n_nodes ? points[i] : -1;
        lsc::PairOpts o{(flags & 1) != 0, (flags & 2) != 0, (flags & 4) != 0, (flags & 8) != 0,
                        false, false, false, false};
        o.network = true;
        lsc::NetPairs net;
        net.emit = emit, net.ctx = ctx;
        lsc::pairwise(G, DG, nullptr, pts, excluded, o, nullptr, nullptr, resistances, nullptr, nullptr, cx, &net);
        int64_t k = 0;
        if (o.maps()) {
            std::copy(net.cum.begin(), net.cum.end(), cum);
            for (lsc::i32 c = 0; c < G.n; ++c)
                for (lsc::i64 e = G.ptr[c]; e < G.ptr[c + 1]; ++e)
                    if (G.adj[e] < c) {
                        if (k >= n_edges) throw std::logic_error("more branches than edges");
                        cum_branch_r[k] = G.adj[e], cum_branch_c[k] = c, cum_branch_current[k] = net.cum_branch[e];
                        ++k;
                    }
        }
        *n_cum_branches = k;
        *n_components = G.ncomp;
        lsc::fill_stats(cx, stats);
    });
}

int lsc_circuitscape_multi_advanced(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                                    int avg_resistances, int four_neighbors, int policy, int64_t nprob,
                                    const int64_t* sptr, const int64_t* scells, const double* svals,
                                    const int64_t* gptr, const int64_t* gcells, const double* gvals,
                                    const int64_t* check, const lsc_solve_options* options, int32_t* nodemap,
                                    double* voltage, double* current, uint8_t* active, int32_t* solved,
                                    lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        if (policy < 0 || policy > 3) throw std::invalid_argument("remove_src_or_gnd must be keepall, rmvsrc, rmvgnd or rmvall");
        lsc::Ctx cx = lsc::make_ctx(options);
        lsc::RGraph G = lsc::raster_graph(gmap, polymap, nx, ny, avg_resistances != 0, four_neighbors != 0);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        std::vector<lsc::Problem> probs;
        std::vector<i32> only;
        lsc::problems_from_cells(G, nprob, sptr, scells, svals, gptr, gcells, gvals, check, probs, only);
        auto sols = lsc::advanced_nodes(G, DG, probs, only, policy, true, cx);
        const i64 n = nx * ny;
        for (i64 c = 0; c < n; ++c) nodemap[c] = G.nodemap[c];
        for (int64_t p = 0; p < nprob; ++p) {
            std::vector<std::uint8_t> act(G.n, 0);
            for (i32 v : sols[p].active) act[v] = 1;
            for (i64 c = 0; c < n; ++c) {
                const i32 v = G.nodemap[c];
                voltage[p * n + c] = v >= 0 ? sols[p].voltage[v] : 0.0;
                current[p * n + c] = v >= 0 ? sols[p].current[v] : 0.0;
                active[p * n + c] = v >= 0 ? act[v] : 0;
            }
            solved[p] = sols[p].solved;
        }
        lsc::fill_stats(cx, stats);
    });
}

int lsc_circuitscape_multi_advanced_sum(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                                        int avg_resistances, int four_neighbors, int policy, int64_t nprob,
                                        const int64_t* sptr, const int64_t* scells, const double* svals,
                                        const int64_t* gptr, const int64_t* gcells, const double* gvals,
                                        const int64_t* check, int focal_zero, const lsc_solve_options* options,
                                        int32_t* nodemap, double* cum, double* max, double* check_voltage,
                                        int32_t* solved, lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        if (policy < 0 || policy > 3) throw std::invalid_argument("remove_src_or_gnd must be keepall, rmvsrc, rmvgnd or rmvall");
        if (!cum) throw std::invalid_argument("the cumulative map is required");
        lsc::Ctx cx = lsc::make_ctx(options);
        const auto h0 = std::chrono::steady_clock::now();
        lsc::progress("multi-advanced: raster graph of %lld x %lld cells", static_cast<long long>(nx),
                      static_cast<long long>(ny));
        lsc::RGraph G = lsc::raster_graph(gmap, polymap, nx, ny, avg_resistances != 0, four_neighbors != 0);
        lsc::progress("multi-advanced: %d nodes, %d components; upload", G.n, G.ncomp);
        lsc::DevGraph DG;
        DG.upload(G, cx.s);
        std::vector<lsc::Problem> probs;
        std::vector<i32> only;
        lsc::progress("multi-advanced: %lld problems from their cells", static_cast<long long>(nprob));
        lsc::problems_from_cells(G, nprob, sptr, scells, svals, gptr, gcells, gvals, check, probs, only);
        std::vector<i32> cnode(nprob, -1);
        for (int64_t p = 0; p < nprob; ++p) cnode[p] = only[p] >= 0 ? only[p] : -1;
        lsc::SumSink sink(G, cx.s, focal_zero != 0, cum, max, cnode);
        lsc::progress("multi-advanced: maps on the device; problems resolved");
        cx.st.ms_fill += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
        const std::vector<char> ok = lsc::advanced_many(G, DG, probs, only, policy, true, cx, sink);
        const i64 n = nx * ny;
        if (max && nprob > 0 && n > 0)
            hipLaunchKernelGGL(lsc::k_max0, dim3(lsc::nblk(n, 256)), dim3(256), 0, cx.s, n, sink.mx.p);
        if (n > 0) {
            LSC_HIP(hipMemcpyAsync(cum, sink.cum.p, n * sizeof(double), hipMemcpyDeviceToHost, cx.s));
            if (max) LSC_HIP(hipMemcpyAsync(max, sink.mx.p, n * sizeof(double), hipMemcpyDeviceToHost, cx.s));
        }
        LSC_HIP(hipStreamSynchronize(cx.s));
        lsc::progress("multi-advanced: maps read back");
        for (i64 c = 0; c < n; ++c) nodemap[c] = G.nodemap[c];
        for (int64_t p = 0; p < nprob; ++p) {
            check_voltage[p] = sink.vcheck[p];
            solved[p] = ok[p];
        }
        lsc::fill_stats(cx, stats);
    });
}

int lsc_grid_circuit_create(const double* resistance, int64_t nx, int64_t ny, int four_neighbors,
                            const lsc_solve_options* options, lsc_grid_circuit** out) {
    return lsc::guarded([&] {
        if (!out || (nx * ny > 0 && !resistance)) throw std::invalid_argument("null argument");
        if (nx < 0 || ny < 0) throw std::invalid_argument("negative raster size");
        auto C = std::make_unique<lsc::GridCircuit>();
        C->cx = lsc::make_ctx(options);
        C->nx = nx, C->ny = ny, C->four = four_neighbors;
        C->valid.resize(static_cast<std::size_t>(nx * ny));
        for (i64 c = 0; c < nx * ny; ++c) C->valid[c] = lsc::valid_resistance(resistance[c]);
        lsc::grid_factor(*C, resistance, false);
        *out = reinterpret_cast<lsc_grid_circuit*>(C.release());
    });
}

int lsc_grid_circuit_update(lsc_grid_circuit* h, const double* resistance) {
    return lsc::guarded([&] {
        if (!h || !resistance) throw std::invalid_argument("null argument");
        auto& C = *reinterpret_cast<lsc::GridCircuit*>(h);
        for (i64 c = 0; c < C.nx * C.ny; ++c)
            if (static_cast<bool>(C.valid[c]) != lsc::valid_resistance(resistance[c]))
                throw std::invalid_argument("the new resistance must keep the circuit's cells (finite r > 0)");
        lsc::grid_factor(C, resistance, true);
    });
}

int lsc_grid_circuit_solve(lsc_grid_circuit* h, int k, const double* rhs_cells, double* out_cells) {
    return lsc::guarded([&] { grid_solve(*reinterpret_cast<lsc::GridCircuit*>(h), k, rhs_cells, out_cells); });
}

int lsc_grid_circuit_adjoint(lsc_grid_circuit* h, int k, const double* dJdV_cells, const double* V_cells,
                             double* grad_r) {
    return lsc::guarded([&] {
        auto& C = *reinterpret_cast<lsc::GridCircuit*>(h);
        const i64 n = C.nx * C.ny;
        lsc::DevBuf<double> lam(static_cast<std::size_t>(k) * n);
        grid_solve(C, k, dJdV_cells, lam.p);
        hipLaunchKernelGGL(lsc::k_adjoint_grad, dim3(lsc::nblk(n, 256)), dim3(256), 0, C.cx.s, C.nx, C.ny, C.four, k,
                           C.r.p, V_cells, lam.p, grad_r);
        LSC_HIP(hipStreamSynchronize(C.cx.s));
    });
}

int lsc_grid_circuit_stats(lsc_grid_circuit* h, lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        auto& C = *reinterpret_cast<lsc::GridCircuit*>(h);
        lsc::Ctx cx = C.cx;
        for (const auto& S : C.sys) lsc::tally(cx, *S);
        lsc::fill_stats(cx, stats);
    });
}

void lsc_grid_circuit_destroy(lsc_grid_circuit* h) { delete reinterpret_cast<lsc::GridCircuit*>(h); }

int lsc_effective_resistance_grad(const double* resistance, int64_t nx, int64_t ny, int four_neighbors,
                                  const int64_t* src_cells, const int64_t* gnd_cells, int64_t npairs,
                                  const lsc_solve_options* options, double* J, double* R, double* grad,
                                  lsc_solve_stats* stats) {
    return lsc::guarded([&] {
        const i64 n = nx * ny;
        lsc_grid_circuit* h = nullptr;
        const int created = lsc_grid_circuit_create(resistance, nx, ny, four_neighbors, options, &h);
        if (created == 1) throw std::invalid_argument(lsc_last_error());
        if (created != 0) throw lsc::HipError(lsc_last_error());
        std::unique_ptr<lsc::GridCircuit> C(reinterpret_cast<lsc::GridCircuit*>(h));
        for (int64_t k = 0; k < npairs; ++k) {
            if (src_cells[k] < 0 || src_cells[k] >= n || gnd_cells[k] < 0 || gnd_cells[k] >= n)
                throw std::invalid_argument("pair cell out of range");
            const i32 a = C->g.nodemap[src_cells[k]], b = C->g.nodemap[gnd_cells[k]];
            if (a < 0 || b < 0 || C->g.comp[a] != C->g.comp[b])
                throw std::invalid_argument("a pair is not in one connected component");
        }
        // dJ/dv of J = sum_k R_k, R_k = v_s - v_t for the RHS e_s - e_t
        std::vector<double> rhs(static_cast<std::size_t>(npairs) * n, 0.0);
        for (int64_t k = 0; k < npairs; ++k) {
            rhs[k * n + src_cells[k]] += 1.0;
            rhs[k * n + gnd_cells[k]] -= 1.0;
        }
        lsc::DevBuf<double> drhs(rhs.size() ? rhs.size() : 1), dV(rhs.size() ? rhs.size() : 1), dg(n ? n : 1);
        drhs.upload(rhs.data(), rhs.size(), C->cx.s);
        grid_solve(*C, static_cast<int>(npairs), drhs.p, dV.p);
        std::vector<double> V(rhs.size());
        LSC_HIP(hipMemcpyAsync(V.data(), dV.p, V.size() * sizeof(double), hipMemcpyDeviceToHost, C->cx.s));
        LSC_HIP(hipStreamSynchronize(C->cx.s));
        double tot = 0.0;
        for (int64_t k = 0; k < npairs; ++k) {
            R[k] = V[k * n + src_cells[k]] - V[k * n + gnd_cells[k]];
            tot += R[k];
        }
        *J = tot;
        // lambda_k = v_k (J is a sum of quadratic forms); the kernel takes both
        hipLaunchKernelGGL(lsc::k_adjoint_grad, dim3(lsc::nblk(n, 256)), dim3(256), 0, C->cx.s, nx, ny,
                           four_neighbors, static_cast<int>(npairs), C->r.p, dV.p, dV.p, dg.p);
        LSC_HIP(hipMemcpyAsync(grad, dg.p, n * sizeof(double), hipMemcpyDeviceToHost, C->cx.s));
        LSC_HIP(hipStreamSynchronize(C->cx.s));
        lsc::Ctx cx = C->cx;
        for (const auto& S : C->sys) lsc::tally(cx, *S);
        lsc::fill_stats(cx, stats);
    });
}

int lsc_circuitscape_analyze(const double* gmap, const int64_t* polymap, int64_t nx, int64_t ny,
                             int avg_resistances, int four_neighbors, lsc_analysis* out) {
    return lsc::guarded([&] {
        if (nx < 0 || ny < 0) throw std::invalid_argument("negative raster size");
        const auto t0 = std::chrono::steady_clock::now();
        lsc::RGraph G = lsc::raster_graph(gmap, polymap, nx, ny, avg_resistances != 0, four_neighbors != 0);
        lsc::analyze(G, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                     out);
    });
}

int lsc_network_analyze(const int32_t* a, const int32_t* b, const double* w, int64_t n_edges, int32_t n_nodes,
                        lsc_analysis* out) {
    return lsc::guarded([&] {
        if (n_nodes <= 0 || n_edges < 0) throw std::invalid_argument("a network needs nodes");
        const auto t0 = std::chrono::steady_clock::now();
        lsc::RGraph G = lsc::network_graph(a, b, w, n_edges, n_nodes);
        lsc::analyze(G, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                     out);
    });
}

}  // extern "C"
