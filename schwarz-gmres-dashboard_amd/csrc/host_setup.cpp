// Host-side setup: partition, overlap sets, extraction, level schedules.
#include "host_setup.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <functional>
#include <numeric>
#include <memory_resource>
#include <set>
#include <stdexcept>
#include <thread>
#include <utility>

namespace sg {

namespace {

// Undirected weighted graph in CSR form.
struct Graph {
    int nv = 0;
    std::vector<int> xadj, adj, vw;
    std::vector<long long> ew;
    long long total_vw() const {
        long long t = 0;
        for (int w : vw) t += w;
        return t;
    }
};

// Threads this thread may use: the hardware's (at most 16) at the top; the
// recursive bisection hands half of its share to each half it runs concurrently.
int hw_threads() {
    return (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}
thread_local int tl_share = 0;

// Runs f(lo, hi) over [0, n) in contiguous chunks on this thread's share of the
// cores (results must not depend on the chunking).
void parallel_rows(int n, const std::function<void(int, int)>& f) {
    const int nt = tl_share > 0 ? tl_share : hw_threads();
    if (nt == 1 || n < 200000) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
        th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

Graph graph_from_csr(const HostCsr& A, bool weighted) {
    // symmetrise the pattern, drop self loops and duplicates; with weighted,
    // edge (i,j) weighs max(|a_ij|, |a_ji|) in units of the smallest nonzero
    // off-diagonal magnitude (at least 1).  Vertex i's neighbours are the
    // union of row i and column i of A, ascending, each with the largest
    // weight of its occurrences: row i (sorted) merged with the transpose's
    // row i, which a counting pass builds in ascending order.
    int n = A.n;
    double amin = 0.0;
    if (weighted) {  // a minimum: the same whatever the chunking
        const int nt = hw_threads();
        std::vector<double> part_min(nt, 1e300);
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
                double m = 1e300;
                for (int i = lo; i < hi; ++i)
                    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
                        if (A.ci[k] != i && A.v[k] != 0.0) m = std::min(m, std::fabs(A.v[k]));
                part_min[t] = m;
            });
        for (auto& x : th) x.join();
        amin = *std::min_element(part_min.begin(), part_min.end());
        if (!(amin < 1e300)) weighted = false;
    }
    std::vector<long long> wk;
    if (weighted) {
        wk.resize(A.rp[n]);
        parallel_rows(n, [&](int lo, int hi) {
            for (int k = A.rp[lo]; k < A.rp[hi]; ++k) {
                const double q = std::fabs(A.v[k]) / amin;
                wk[k] = q >= 1e12 ? 1000000000000LL : std::max(1LL, (long long)std::llround(q));
            }
        });
    }
    auto wt = [&](int k) -> long long { return weighted ? wk[k] : 1; };
    auto valid = [&](int i, int j) { return j != i && j >= 0 && j < n; };
    // A symmetric matrix (pattern and weights, rows strictly ascending) is its
    // own transpose: the union of row i and column i is row i, so the
    // transpose is not built (the graph is the same).
    bool symmetric = true;
    {
        std::vector<char> ok(hw_threads(), 1);
        std::vector<std::thread> th;
        const int nt = (int)ok.size();
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
                for (int i = lo; i < hi && ok[t]; ++i)
                    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                        const int j = A.ci[k];
                        if (k > A.rp[i] && A.ci[k - 1] >= j) { ok[t] = 0; break; }
                        if (!valid(i, j)) continue;
                        const int* b = A.ci.data() + A.rp[j];
                        const int* e = A.ci.data() + A.rp[j + 1];
                        const int* f = std::lower_bound(b, e, i);
                        if (f == e || *f != i || wt((int)(f - A.ci.data())) != wt(k)) {
                            ok[t] = 0;
                            break;
                        }
                    }
            });
        for (auto& x : th) x.join();
        for (char c : ok) symmetric = symmetric && c;
    }
    // transpose pattern: tr row j lists every i with a valid entry (i, j), ascending i
    std::vector<int> tp(n + 1, 0);
    std::vector<int> ti;
    std::vector<long long> tw;
    if (!symmetric) {
        for (int i = 0; i < n; ++i)
            for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
                if (valid(i, A.ci[k])) tp[A.ci[k] + 1]++;
        for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
        ti.resize(tp[n]);
        tw.resize(tp[n]);
        std::vector<int> fill(tp.begin(), tp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int j = A.ci[k];
                if (!valid(i, j)) continue;
                ti[fill[j]] = i;
                tw[fill[j]++] = wt(k);
            }
    }
    // row i's own entries, sorted by column with the largest weight first
    auto own_row = [&](int i, std::vector<std::pair<int, long long>>& r) {
        r.clear();
        bool sorted = true;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (valid(i, A.ci[k])) {
                if (!r.empty() && r.back().first >= A.ci[k]) sorted = false;
                r.push_back({A.ci[k], wt(k)});
            }
        if (sorted) return;  // strictly ascending already: nothing to order or merge
        std::sort(r.begin(), r.end(), [](const std::pair<int, long long>& x,
                                         const std::pair<int, long long>& y) {
            return x.first != y.first ? x.first < y.first : x.second > y.second;
        });
    };
    // merged neighbour list of i (emit(b, w) in ascending b, one per neighbour)
    auto merge = [&](int i, std::vector<std::pair<int, long long>>& r, auto&& emit) {
        own_row(i, r);
        size_t a = 0;
        int b = tp[i];
        const int be = tp[i + 1];
        while (a < r.size() || b < be) {
            int c;
            long long w = -1;
            if (b >= be || (a < r.size() && r[a].first <= ti[b])) c = r[a].first;
            else c = ti[b];
            while (a < r.size() && r[a].first == c) w = std::max(w, r[a++].second);
            while (b < be && ti[b] == c) w = std::max(w, tw[b++]);
            emit(c, w);
        }
    };
    Graph g;
    g.nv = n;
    g.xadj.assign(n + 1, 0);
    parallel_rows(n, [&](int lo, int hi) {
        std::vector<std::pair<int, long long>> r;
        for (int i = lo; i < hi; ++i) {
            int cnt = 0;
            merge(i, r, [&](int, long long) { ++cnt; });
            g.xadj[i + 1] = cnt;
        }
    });
    for (int i = 0; i < n; ++i) g.xadj[i + 1] += g.xadj[i];
    g.adj.resize(g.xadj[n]);
    g.ew.resize(g.xadj[n]);
    parallel_rows(n, [&](int lo, int hi) {
        std::vector<std::pair<int, long long>> r;
        for (int i = lo; i < hi; ++i) {
            int pos = g.xadj[i];
            merge(i, r, [&](int c, long long w) {
                g.adj[pos] = c;
                g.ew[pos++] = w;
            });
        }
    });
    g.vw.assign(n, 1);
    return g;
}

// Heavy-edge matching coarsening.  Returns coarse graph and fine->coarse map.
bool coarsen(const Graph& g, Graph& c, std::vector<int>& cmap) {
    int n = g.nv;
    std::vector<int> match(n, -1);
    // visit in order of increasing degree (ties by index): deterministic
    // (a stable counting sort on the degree: ties keep index order)
    int maxdeg = 0;
    for (int v = 0; v < n; ++v) maxdeg = std::max(maxdeg, g.xadj[v + 1] - g.xadj[v]);
    std::vector<int> start(maxdeg + 2, 0), order(n);
    for (int v = 0; v < n; ++v) start[g.xadj[v + 1] - g.xadj[v] + 1]++;
    for (int d = 0; d <= maxdeg; ++d) start[d + 1] += start[d];
    for (int v = 0; v < n; ++v) order[start[g.xadj[v + 1] - g.xadj[v]]++] = v;
    long long maxvw = std::max<long long>(1, g.total_vw() / 100);
    for (int v : order) {
        if (match[v] >= 0) continue;
        int best = -1;
        long long bw = -1;
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
            int u = g.adj[k];
            if (match[u] >= 0 || u == v) continue;
            if (g.vw[u] + g.vw[v] > maxvw) continue;
            if (g.ew[k] > bw) { bw = g.ew[k]; best = u; }
        }
        if (best >= 0) { match[v] = best; match[best] = v; }
        else match[v] = v;
    }
    cmap.assign(n, -1);
    int nc = 0;
    for (int v = 0; v < n; ++v) {
        if (cmap[v] >= 0) continue;
        cmap[v] = nc;
        if (match[v] != v) cmap[match[v]] = nc;
        ++nc;
    }
    if (nc > 0.95 * n) return false;
    c.nv = nc;
    c.vw.assign(nc, 0);
    for (int v = 0; v < n; ++v) c.vw[cmap[v]] += g.vw[v];
    // members of each coarse vertex in ascending fine order
    std::vector<int> first(nc, -1), second(nc, -1);
    for (int v = 0; v < n; ++v) {
        int cv = cmap[v];
        if (first[cv] < 0) first[cv] = v; else second[cv] = v;
    }
    // coarse adjacency: the members' neighbours in order of first appearance,
    // weights summed.  Each coarse row depends only on its members' rows, so
    // with several threads rows are counted, then filled, in parallel chunks
    // (the same graph as the single pass).
    auto row_of = [&](int cv, std::vector<std::pair<int, long long>>& r,
                      std::vector<std::pair<int, int>>& srt) {
        r.clear();
        for (int m = 0; m < 2; ++m) {
            const int v = m == 0 ? first[cv] : second[cv];
            if (v < 0) continue;
            for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
                const int cu = cmap[g.adj[k]];
                if (cu != cv) r.push_back({cu, g.ew[k]});
            }
        }
        if (r.size() <= 64) {  // short rows: merge in place, first appearance kept
            size_t w = 0;
            for (size_t a = 0; a < r.size(); ++a) {
                size_t q = 0;
                while (q < w && r[q].first != r[a].first) ++q;
                if (q < w) r[q].second += r[a].second;
                else r[w++] = r[a];
            }
            r.resize(w);
            return;
        }
        // long rows: sort (neighbour, position), sum runs, back to first-appearance order
        srt.resize(r.size());
        for (size_t a = 0; a < r.size(); ++a) srt[a] = {r[a].first, (int)a};
        std::sort(srt.begin(), srt.end());
        std::vector<std::pair<int, long long>> u;  // (first position, weight)
        std::vector<int> who;
        for (size_t a = 0; a < srt.size();) {
            size_t b = a;
            long long w = 0;
            while (b < srt.size() && srt[b].first == srt[a].first) w += r[srt[b++].second].second;
            u.push_back({srt[a].second, w});
            who.push_back(srt[a].first);
            a = b;
        }
        std::vector<int> ord(u.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int x, int y) { return u[x].first < u[y].first; });
        r.resize(ord.size());
        for (size_t a = 0; a < ord.size(); ++a) r[a] = {who[ord[a]], u[ord[a]].second};
    };
    c.xadj.assign(nc + 1, 0);
    const int nt = tl_share > 0 ? tl_share : hw_threads();
    if (nt == 1 || nc < 200000) {  // one thread: a single pass with a mark array
        c.adj.clear();
        c.ew.clear();
        c.adj.reserve(g.adj.size() / 2);
        c.ew.reserve(g.adj.size() / 2);
        std::vector<int> mark(nc, -1), slot(nc, 0);
        for (int cv = 0; cv < nc; ++cv) {
            for (int m = 0; m < 2; ++m) {
                const int v = m == 0 ? first[cv] : second[cv];
                if (v < 0) continue;
                for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
                    const int cu = cmap[g.adj[k]];
                    if (cu == cv) continue;
                    if (mark[cu] != cv) {
                        mark[cu] = cv;
                        slot[cu] = (int)c.adj.size();
                        c.adj.push_back(cu);
                        c.ew.push_back(g.ew[k]);
                    } else {
                        c.ew[slot[cu]] += g.ew[k];
                    }
                }
            }
            c.xadj[cv + 1] = (int)c.adj.size();
        }
        return true;
    }
    parallel_rows(nc, [&](int lo, int hi) {
        std::vector<std::pair<int, long long>> r;
        std::vector<std::pair<int, int>> srt;
        for (int cv = lo; cv < hi; ++cv) {
            row_of(cv, r, srt);
            c.xadj[cv + 1] = (int)r.size();
        }
    });
    for (int cv = 0; cv < nc; ++cv) c.xadj[cv + 1] += c.xadj[cv];
    c.adj.resize(c.xadj[nc]);
    c.ew.resize(c.xadj[nc]);
    parallel_rows(nc, [&](int lo, int hi) {
        std::vector<std::pair<int, long long>> r;
        std::vector<std::pair<int, int>> srt;
        for (int cv = lo; cv < hi; ++cv) {
            row_of(cv, r, srt);
            int pos = c.xadj[cv];
            for (const auto& e : r) {
                c.adj[pos] = e.first;
                c.ew[pos++] = e.second;
            }
        }
    });
    return true;
}

long long cut_of(const Graph& g, const std::vector<int>& side) {
    long long cut = 0;
    for (int v = 0; v < g.nv; ++v)
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
            if (side[g.adj[k]] != side[v]) cut += g.ew[k];
    return cut / 2;
}

// Boundary Fiduccia-Mattheyses refinement of a bisection with balance window
// [lo, hi] on the weight of side 0.
void fm_refine(const Graph& g, std::vector<int>& side, long long lo, long long hi, int passes) {
    int n = g.nv;
    long long w0 = 0;
    for (int v = 0; v < n; ++v) if (side[v] == 0) w0 += g.vw[v];
    long long minvw = 1, maxvw = 1;
    if (n > 0) {
        const auto mm = std::minmax_element(g.vw.begin(), g.vw.end());
        minvw = *mm.first;
        maxvw = *mm.second;
    }
    std::vector<long long> gain(n);
    std::vector<int> locked(n);
    auto compute_gain = [&](int v) {
        long long ext = 0, in = 0;
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
            (side[g.adj[k]] != side[v] ? ext : in) += g.ew[k];
        return ext - in;
    };
    auto boundary = [&](int v) {
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
            if (side[g.adj[k]] != side[v]) return true;
        return false;
    };
    // bnd[v] and gain[v] (for boundary v) hold at the start of every pass:
    // the first pass computes them on all threads; a pass changes sides only
    // of the vertices it moved, so before the next one they are recomputed
    // for those and their neighbours only, and the boundary list is updated.
    std::vector<char> bnd(n), inlist(n, 0), inq(n, 0);
    std::vector<int> blist, stamp(n, -1);
    parallel_rows(n, [&](int lo, int hi) {
        for (int v = lo; v < hi; ++v) {
            bnd[v] = boundary(v);
            if (bnd[v]) gain[v] = compute_gain(v);
        }
    });
    for (int v = 0; v < n; ++v)
        if (bnd[v]) {
            blist.push_back(v);
            inlist[v] = 1;
        }
    std::vector<int> moved;
    for (int pass = 0; pass < passes; ++pass) {
        if (pass > 0) {
            for (int v : moved) locked[v] = 0;
            std::vector<int> touched;
            auto touch = [&](int v) {
                if (stamp[v] == pass) return;
                stamp[v] = pass;
                bnd[v] = boundary(v);
                if (bnd[v]) gain[v] = compute_gain(v);
                if (bnd[v] && !inlist[v]) touched.push_back(v);
            };
            for (int v : moved) {
                touch(v);
                for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) touch(g.adj[k]);
            }
            size_t keep = 0;
            for (int v : blist)
                if (bnd[v]) blist[keep++] = v;
                else inlist[v] = 0;
            blist.resize(keep);
            for (int v : touched) {
                blist.push_back(v);
                inlist[v] = 1;
            }
        }
        // the queues are ordered sets (the order of insertion does not
        // matter): filled from sorted keys with end hints, nodes from a pool
        std::pmr::unsynchronized_pool_resource pool;
        std::pmr::set<std::pair<long long, int>> q[2] = {
            std::pmr::set<std::pair<long long, int>>(&pool),
            std::pmr::set<std::pair<long long, int>>(&pool)};  // (-gain, v)
        {
            std::vector<std::pair<long long, int>> keys[2];
            for (int v : blist) keys[side[v]].push_back({-gain[v], v});
            for (int s = 0; s < 2; ++s) {
                std::sort(keys[s].begin(), keys[s].end());
                for (const auto& k : keys[s]) {
                    q[s].insert(q[s].end(), k);
                    inq[k.second] = 1;
                }
            }
        }
        moved.clear();
        long long cur = 0, best = 0;
        size_t best_len = 0;
        long long best_imb = std::llabs(w0 - (lo + hi) / 2);
        bool best_ok = w0 >= lo && w0 <= hi;
        int since_best = 0;
        const int limit = std::max(50, std::min(n / 20 + 1, 400));
        while (since_best < limit) {
            // candidate from each side that keeps the balance window
            int pick = -1;
            int pside = -1;
            for (int s = 0; s < 2; ++s) {
                // feasible: the move keeps side 0's weight in the window, or
                // moves it toward the window from outside.  With w0 fixed that
                // is an interval [L, U] of the vertex's weight (weights >= 1).
                long long L, U;
                if (s == 0) {
                    U = w0 - lo;
                    L = w0 > hi ? 1 : w0 - hi;
                } else {
                    U = hi - w0;
                    L = w0 < lo ? 1 : lo - w0;
                }
                if (U < L || U < minvw || L > maxvw) continue;  // no vertex fits
                const bool all_fit = L <= minvw && maxvw <= U;   // every vertex fits
                for (auto it = q[s].begin(); it != q[s].end(); ++it) {
                    int v = it->second;
                    if (!all_fit && (g.vw[v] < L || g.vw[v] > U)) continue;
                    long long nw0 = s == 0 ? w0 - g.vw[v] : w0 + g.vw[v];
                    if (pick < 0 || gain[v] > gain[pick] ||
                        (gain[v] == gain[pick] && std::llabs(nw0 - (lo + hi) / 2) <
                                                      std::llabs((pside == 0 ? w0 - g.vw[pick] : w0 + g.vw[pick]) - (lo + hi) / 2))) {
                        pick = v;
                        pside = s;
                    }
                    break;  // only the best-gain feasible-first entry per side
                }
            }
            if (pick < 0) break;
            q[pside].erase({-gain[pick], pick});
            inq[pick] = 0;
            cur -= gain[pick];
            w0 += pside == 0 ? -g.vw[pick] : g.vw[pick];
            side[pick] ^= 1;
            locked[pick] = 1;
            moved.push_back(pick);
            for (int k = g.xadj[pick]; k < g.xadj[pick + 1]; ++k) {
                const int u = g.adj[k];
                if (locked[u]) continue;
                if (inq[u]) q[side[u]].erase({-gain[u], u});
                // gain and boundary in one pass (edge weights are >= 1, so
                // u is on the boundary exactly when ext > 0)
                long long ext = 0, in = 0;
                for (int kk = g.xadj[u]; kk < g.xadj[u + 1]; ++kk)
                    (side[g.adj[kk]] != side[u] ? ext : in) += g.ew[kk];
                gain[u] = ext - in;
                inq[u] = ext > 0;
                if (inq[u]) q[side[u]].insert({-gain[u], u});
            }
            long long imb = std::llabs(w0 - (lo + hi) / 2);
            bool ok_now = w0 >= lo && w0 <= hi;
            if ((ok_now && !best_ok) || (ok_now == best_ok && (cur < best || (cur == best && imb < best_imb)))) {
                best_ok = ok_now;
                best = cur;
                best_len = moved.size();
                best_imb = imb;
                since_best = 0;
            } else {
                ++since_best;
            }
        }
        for (int s2 = 0; s2 < 2; ++s2)
            for (const auto& e : q[s2]) inq[e.second] = 0;
        // roll back past the best prefix
        for (size_t i = moved.size(); i > best_len; --i) {
            int v = moved[i - 1];
            w0 += side[v] == 0 ? -g.vw[v] : g.vw[v];
            side[v] ^= 1;
        }
        if (best_len == 0) break;
    }
}

// Greedy graph growing from a seed; grows side 0 until its weight >= target.
std::vector<int> grow(const Graph& g, int seed, long long target) {
    int n = g.nv;
    std::vector<int> side(n, 1), inq(n, 0);
    std::vector<long long> gain(n, 0);
    std::set<std::pair<long long, int>> q;
    long long w0 = 0;
    auto push = [&](int v) {
        long long ext = 0, in = 0;
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
            (side[g.adj[k]] == 0 ? in : ext) += g.ew[k];
        gain[v] = in - ext;
        q.insert({-gain[v], v});
        inq[v] = 1;
    };
    push(seed);
    int next_unreached = 0;
    while (w0 < target) {
        if (q.empty()) {  // disconnected: restart from the next side-1 vertex
            while (next_unreached < n && side[next_unreached] == 0) ++next_unreached;
            if (next_unreached >= n) break;
            push(next_unreached);
        }
        int v = q.begin()->second;
        q.erase(q.begin());
        inq[v] = 0;
        side[v] = 0;
        w0 += g.vw[v];
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
            int u = g.adj[k];
            if (side[u] == 0) continue;
            if (inq[u]) q.erase({-gain[u], u});
            push(u);
        }
    }
    return side;
}

int bfs_far(const Graph& g, int src) {
    std::vector<int> d(g.nv, -1);
    std::vector<int> qu;
    qu.push_back(src);
    d[src] = 0;
    int last = src;
    for (size_t h = 0; h < qu.size(); ++h) {
        int v = qu[h];
        last = v;
        for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
            if (d[g.adj[k]] < 0) { d[g.adj[k]] = d[v] + 1; qu.push_back(g.adj[k]); }
    }
    return last;
}

std::vector<int> initial_bisect(const Graph& g, long long target, long long lo, long long hi) {
    std::vector<int> seeds;
    int a = bfs_far(g, 0);
    int b = bfs_far(g, a);
    seeds.push_back(a);
    seeds.push_back(b);
    for (int t = 1; t <= 4 && g.nv > 8; ++t) seeds.push_back((int)((long long)g.nv * t / 5));
    std::vector<int> best;
    long long bestcut = -1;
    for (int s : seeds) {
        std::vector<int> side = grow(g, s, target);
        fm_refine(g, side, lo, hi, 4);
        long long c = cut_of(g, side);
        if (bestcut < 0 || c < bestcut) { bestcut = c; best = side; }
    }
    return best;
}

// Multilevel bisection: side[v] in {0,1}, side 0 weight ~ frac * total.
std::vector<int> bisect(const Graph& g, double frac, double eps) {
    long long tot = g.total_vw();
    long long target = (long long)std::llround(frac * (double)tot);
    // level 0 is g itself (not copied); coarser levels are owned here
    std::vector<Graph> coarse;
    std::vector<std::vector<int>> maps;
    coarse.reserve(64);
    auto level = [&](int l) -> const Graph& { return l == 0 ? g : coarse[l - 1]; };
    int nlev = 1;
    while (level(nlev - 1).nv > 160) {
        Graph c;
        std::vector<int> cm;
        if (!coarsen(level(nlev - 1), c, cm)) break;
        maps.push_back(std::move(cm));
        coarse.push_back(std::move(c));
        ++nlev;
    }
    // side s may exceed its target share t_s by eps * t_s (and by one coarse
    // vertex while the graph is coarse)
    auto window = [&](const Graph& h) {
        long long maxv = 1;
        for (int w : h.vw) maxv = std::max<long long>(maxv, w);
        const double t0 = (double)target, t1 = (double)(tot - target);
        long long up0 = std::max<long long>((long long)(eps * t0), maxv);
        long long up1 = std::max<long long>((long long)(eps * t1), maxv);
        return std::make_pair(target - up1, target + up0);
    };
    auto w = window(level(nlev - 1));
    std::vector<int> side = initial_bisect(level(nlev - 1), target, w.first, w.second);
    for (int l = nlev - 2; l >= 0; --l) {
        const Graph& gl = level(l);
        std::vector<int> fine(gl.nv);
        parallel_rows(gl.nv, [&](int lo, int hi) {
            for (int v = lo; v < hi; ++v) fine[v] = side[maps[l][v]];
        });
        side.swap(fine);
        auto wl = window(gl);
        fm_refine(gl, side, wl.first, wl.second, l == 0 ? 6 : 3);
    }
    return side;
}

Graph induced(const Graph& g, const std::vector<int>& verts, std::vector<int>& local) {
    // verts ascending; local[v] = new id or -1; rows counted, then filled, in parallel
    Graph s;
    s.nv = (int)verts.size();
    s.xadj.assign(s.nv + 1, 0);
    s.vw.resize(s.nv);
    parallel_rows(s.nv, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            int v = verts[i], cnt = 0;
            s.vw[i] = g.vw[v];
            for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) cnt += local[g.adj[k]] >= 0;
            s.xadj[i + 1] = cnt;
        }
    });
    for (int i = 0; i < s.nv; ++i) s.xadj[i + 1] += s.xadj[i];
    s.adj.resize(s.xadj[s.nv]);
    s.ew.resize(s.xadj[s.nv]);
    parallel_rows(s.nv, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            int v = verts[i], pos = s.xadj[i];
            for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
                int u = local[g.adj[k]];
                if (u >= 0) {
                    s.adj[pos] = u;
                    s.ew[pos++] = g.ew[k];
                }
            }
        }
    });
    return s;
}

void recurse(const Graph& g, const std::vector<int>& ids, int nparts, int first, double eps,
             std::vector<int>& part, int depth) {
    if (nparts <= 1 || g.nv == 0) {
        for (int id : ids) part[id] = first;
        return;
    }
    int k0 = nparts / 2, k1 = nparts - k0;
    std::vector<int> side = bisect(g, (double)k0 / nparts, eps);
    std::vector<int> v0, v1, loc0(g.nv, -1), loc1(g.nv, -1);
    for (int v = 0; v < g.nv; ++v) {
        if (side[v] == 0) { loc0[v] = (int)v0.size(); v0.push_back(v); }
        else { loc1[v] = (int)v1.size(); v1.push_back(v); }
    }
    Graph g0 = induced(g, v0, loc0), g1 = induced(g, v1, loc1);
    std::vector<int> ids0(v0.size()), ids1(v1.size());
    for (size_t i = 0; i < v0.size(); ++i) ids0[i] = ids[v0[i]];
    for (size_t i = 0; i < v1.size(); ++i) ids1[i] = ids[v1[i]];
    // the two halves are independent: recurse concurrently while this
    // thread's share of the cores has room for both
    const int share = tl_share > 0 ? tl_share : hw_threads();
    if (share >= 2 && g.nv > 20000) {
        const int half = std::max(1, share / 2);
        std::thread t([&, half] {
            tl_share = half;
            recurse(g0, ids0, k0, first, eps, part, depth + 1);
        });
        const int mine = tl_share;
        tl_share = std::max(1, share - half);
        recurse(g1, ids1, k1, first + k0, eps, part, depth + 1);
        tl_share = mine;
        t.join();
    } else {
        recurse(g0, ids0, k0, first, eps, part, depth + 1);
        recurse(g1, ids1, k1, first + k0, eps, part, depth + 1);
    }
}

}  // namespace

std::vector<int> partition_graph(const HostCsr& A, int nparts, int ufactor, bool weighted) {
    if (nparts <= 1) return std::vector<int>(A.n, 0);
    if (nparts > A.n) throw std::runtime_error("partitioning failed: more parts than rows");
    if (ufactor <= 0) ufactor = 30;
    Graph g = graph_from_csr(A, weighted);
    int depth = (int)std::ceil(std::log2((double)nparts));
    double eps = std::pow(1.0 + ufactor / 1000.0, 1.0 / std::max(1, depth)) - 1.0;
    std::vector<int> part(A.n, -1), ids(A.n);
    std::iota(ids.begin(), ids.end(), 0);
    recurse(g, ids, nparts, 0, eps, part, 0);
    std::vector<int> count(nparts, 0);
    for (int p : part) {
        if (p < 0 || p >= nparts) throw std::runtime_error("partitioning failed");
        count[p]++;
    }
    for (int c : count)
        if (c == 0) throw std::runtime_error("partitioning failed: empty subdomain");
    return part;
}

void build_sets(const HostCsr& A, const std::vector<int>& part, int nparts, int overlap,
                std::vector<std::vector<int>>& owned, std::vector<std::vector<int>>& dofs) {
    int n = A.n;
    owned.assign(nparts, {});
    for (int i = 0; i < n; ++i) owned[part[i]].push_back(i);
    dofs.assign(nparts, {});
    // Subdomains are independent: threads take them from a shared counter,
    // each with its own mark array (stamped with the subdomain), so the sets
    // do not depend on which thread builds which subdomain.
    std::atomic<int> next{0};
    auto work = [&] {
        std::vector<int> mark(std::max(n, A.ncols), -1);
        for (int s = next++; s < nparts; s = next++) {
            std::vector<int> cur = owned[s], frontier = owned[s];
            for (int i : cur) mark[i] = s;
            // Layer l adds every column of every row already in the set that
            // is not yet in it; only the previous layer can contribute new
            // columns.
            for (int layer = 0; layer < overlap; ++layer) {
                std::vector<int> add;
                for (int i : frontier)
                    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                        const int c = A.ci[k];
                        if (mark[c] != s) {
                            mark[c] = s;
                            add.push_back(c);
                        }
                    }
                cur.insert(cur.end(), add.begin(), add.end());
                frontier.swap(add);
                if (frontier.empty()) break;
            }
            std::sort(cur.begin(), cur.end());
            dofs[s] = std::move(cur);
        }
    };
    const int nt = std::min(hw_threads(), std::max(1, nparts));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
}

HostCsr extract_submatrix(const HostCsr& A, const std::vector<int>& idx) {
    int m = (int)idx.size();
    thread_local std::vector<int> inv;
    if ((int)inv.size() < std::max(A.n, A.ncols)) inv.assign(std::max(A.n, A.ncols), -1);
    for (int i = 0; i < m; ++i) inv[idx[i]] = i;
    HostCsr s;
    s.n = m;
    s.ncols = m;
    s.rp.assign(m + 1, 0);
    for (int li = 0; li < m; ++li) {
        int gi = idx[li];
        s.rp[li] = (int)s.ci.size();
        for (int k = A.rp[gi]; k < A.rp[gi + 1]; ++k) {
            int lj = inv[A.ci[k]];
            if (lj >= 0) { s.ci.push_back(lj); s.v.push_back(A.v[k]); }
        }
    }
    s.rp[m] = (int)s.ci.size();
    for (int i = 0; i < m; ++i) inv[idx[i]] = -1;
    return s;
}

std::vector<int> diag_positions(const HostCsr& A) {
    std::vector<int> d(A.n, -1);
    for (int i = 0; i < A.n; ++i)
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i) { d[i] = k; break; }
    return d;
}

std::vector<int> lower_levels(const HostCsr& A) {
    std::vector<int> lev(A.n, 0);
    for (int i = 0; i < A.n; ++i) {
        int l = 0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            int c = A.ci[k];
            if (c >= i) break;  // sorted columns
            l = std::max(l, lev[c] + 1);
        }
        lev[i] = l;
    }
    return lev;
}

std::vector<int> upper_levels(const HostCsr& A) {
    std::vector<int> lev(A.n, 0);
    std::vector<int> d = diag_positions(A);
    for (int i = A.n - 1; i >= 0; --i) {
        int l = 0;
        int k0 = d[i] >= 0 ? d[i] + 1 : A.rp[i + 1];
        for (int k = k0; k < A.rp[i + 1]; ++k) l = std::max(l, lev[A.ci[k]] + 1);
        lev[i] = l;
    }
    return lev;
}

// ILUT(drop_tol, fill_factor).  Row i: scatter A's row into the dense work
// row w (an index joins the list when w[j] is still 0.0), tolerance
// drop_tol * max|w|; sort the list; eliminate with each listed k < i in
// ascending order -- only the indices listed before elimination starts, as the
// reference's range-for over the growing list visits (ilut.h:62) -- dropping
// w[k] below the tolerance, else w[k] /= u_kk and w -= w[k] * (U row k right
// of k), new fill joining the list; sort again; split into L (j < i, kept if
// |w_j| >= tol) and U (j >= i, the diagonal always, others if >= tol), zeros
// skipped; keep the max_fill largest of each by std::partial_sort on |value|
// and re-sort by column (std::sort); u_ii is the U entry in column i, 1e-30 if
// absent or smaller.
void ilut_factor(const HostCsr& A, double drop_tol, int fill_factor, HostCsr& F,
                 std::vector<int>& diag, long long& l_nnz, long long& u_nnz) {
    using Ent = std::pair<int, double>;
    const int n = A.n;
    F = HostCsr{};
    F.n = F.ncols = n;
    F.rp.assign((size_t)n + 1, 0);
    diag.assign(n, -1);
    l_nnz = u_nnz = 0;
    if (n == 0) return;
    const int avg = std::max(1, A.nnz() / n);
    const int max_fill = fill_factor * avg;
    std::vector<std::vector<Ent>> Lr(n), Ur(n);
    std::vector<double> udiag(n);
    std::vector<double> w(n, 0.0);
    std::vector<int> nz;
    auto by_mag = [](const Ent& a, const Ent& b) { return std::abs(a.second) > std::abs(b.second); };
    auto by_col = [](const Ent& a, const Ent& b) { return a.first < b.first; };
    auto keep_largest = [&](std::vector<Ent>& part) {
        if ((int)part.size() > max_fill) {
            std::partial_sort(part.begin(), part.begin() + max_fill, part.end(), by_mag);
            part.resize(max_fill);
            std::sort(part.begin(), part.end(), by_col);
        }
    };
    for (int i = 0; i < n; ++i) {
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const int j = A.ci[p];
            if (w[j] == 0.0) nz.push_back(j);
            w[j] += A.v[p];
        }
        double rmax = 0.0;
        for (int j : nz) rmax = std::max(rmax, std::abs(w[j]));
        const double tol = drop_tol * rmax;
        std::sort(nz.begin(), nz.end());
        const size_t listed = nz.size();
        for (size_t t = 0; t < listed; ++t) {
            const int k = nz[t];
            if (k >= i) break;
            if (std::abs(w[k]) < tol) {
                w[k] = 0.0;
                continue;
            }
            const double ukk = udiag[k];
            if (std::abs(ukk) < 1e-30) {
                w[k] = 0.0;
                continue;
            }
            const double lik = w[k] / ukk;
            w[k] = lik;
            for (const Ent& e : Ur[k]) {
                if (e.first <= k) continue;
                if (w[e.first] == 0.0) nz.push_back(e.first);
                w[e.first] -= lik * e.second;
            }
        }
        std::sort(nz.begin(), nz.end());
        std::vector<Ent> lp, up;
        for (int j : nz) {
            if (std::abs(w[j]) == 0.0) continue;
            if (j < i) {
                if (std::abs(w[j]) >= tol) lp.emplace_back(j, w[j]);
            } else if (j == i || std::abs(w[j]) >= tol) {
                up.emplace_back(j, w[j]);
            }
        }
        keep_largest(lp);
        keep_largest(up);
        double d = 0.0;
        for (const Ent& e : up)
            if (e.first == i) {
                d = e.second;
                break;
            }
        if (std::abs(d) < 1e-30) d = 1e-30;
        udiag[i] = d;
        Lr[i] = std::move(lp);
        Ur[i] = std::move(up);
        for (int j : nz) w[j] = 0.0;
        nz.clear();
    }
    for (int i = 0; i < n; ++i) {
        l_nnz += (long long)Lr[i].size();
        u_nnz += (long long)Ur[i].size();
        for (const Ent& e : Lr[i]) {
            F.ci.push_back(e.first);
            F.v.push_back(e.second);
        }
        diag[i] = (int)F.ci.size();
        F.ci.push_back(i);
        F.v.push_back(udiag[i]);
        for (size_t t = 1; t < Ur[i].size(); ++t) {
            if (Ur[i][t].first <= i)
                throw std::runtime_error("ILUT: repeated diagonal entry in a U row");
            F.ci.push_back(Ur[i][t].first);
            F.v.push_back(Ur[i][t].second);
        }
        if (F.ci.size() > (size_t)INT32_MAX) throw std::runtime_error("ILUT factors exceed 32-bit indexing");
        F.rp[i + 1] = (int)F.ci.size();
    }
}

}  // namespace sg
