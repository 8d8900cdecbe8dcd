// moves_host.hip — the lane-uniform templates of the table-driven random moves (host code).
//
// Restates the network-only parts of the reference's move construction:
//   xloc / xlocc  `xloc_directions` (src/local.jl:122-191): partners within one hop, the angle
//                 change on the ball B = {n, m} that solves L_BB Δθ_B = e_n − e_m with Δθ = 0
//                 outside B (exact global re-routing when B is the whole component), and the
//                 leak into the first layer outside B (requirement change per unit t);
//   xpair         `reroute!` (:44-110): Δθ = L⁺(e_n − e_m) on the component, grounded at its
//                 first node, by a dense Cholesky factor of the grounded Laplacian;
//   cell          the cells of every partition level (E.parts, src/energy.jl:292; LocalModel
//                 :265-283) and the level probabilities ∝ (number of cells)^1 (:282-283).
// The direction of a template is then finished as `finalize!` does (:656-757): angle entries
// below 1e-14 are zero, the lines of the angle entries with |Δf| > 1e-14 are kept, and the
// touched nodes are the nodes of the gas entries, of the kinks and of the kept lines. Lines
// and nodes are listed in increasing index order (the order of the evaluation sums).
#include <math.h>

#include <algorithm>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "esm_internal.h"

namespace esm {

namespace {

// A = G Gᵀ in place (row-major, lower triangle); false if A is not positive definite
bool cholesky(std::vector<double> &A, int n) {
    for (int i = 0; i < n; i++) {
        for (int j = 0; j <= i; j++) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            if (i == j) {
                if (!(s > 0)) return false;
                A[(size_t)i * n + i] = sqrt(s);
            } else {
                A[(size_t)i * n + j] = s / A[(size_t)j * n + j];
            }
        }
        for (int j = i + 1; j < n; j++) A[(size_t)i * n + j] = 0.0;
    }
    return true;
}

// solve G Gᵀ x = b in place
void chol_solve(const std::vector<double> &G, int n, double *b) {
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= G[(size_t)i * n + k] * b[k];
        b[i] = s / G[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < n; k++) s -= G[(size_t)k * n + i] * b[k];
        b[i] = s / G[(size_t)i * n + i];
    }
}

struct Dir {
    std::vector<std::pair<int, double>> th;     // angle entries (node, change per unit t)
    struct Ent {
        int node;
        double c;
        int div;
    };
    std::vector<Ent> en;                         // gas entries
    std::vector<std::pair<int, double>> kk;     // kinks (leak node, requirement change per unit t)
};

struct Net {
    int N = 0, L = 0;
    const int *bus0 = nullptr, *bus1 = nullptr;
    const double *b = nullptr;
    std::vector<std::vector<int>> nodelines;     // lines at each node, in line order
    std::vector<int> comp;                       // component of each node
    std::vector<std::vector<int>> cnodes;        // nodes of each component, increasing
};

// finish a direction (finalize!) and append it as the next template
void emit(MoveTablesHost &mt, const Net &net, const Dir &d) {
    std::vector<double> dth(net.N, 0.0);
    std::vector<int> thn;
    for (const auto &p : d.th) {
        if (fabs(p.second) < 1e-14) continue;
        dth[p.first] = p.second;
        thn.push_back(p.first);
        mt.th_node.push_back(p.first);
        mt.th_val.push_back(p.second);
    }
    std::set<int> lset;
    for (int i : thn)
        for (int e : net.nodelines[i]) lset.insert(e);
    std::vector<int> lines;
    std::vector<double> dfs;
    for (int e : lset) {
        const double df = net.b[e] * (dth[net.bus0[e]] - dth[net.bus1[e]]);
        if (fabs(df) <= 1e-14) continue;
        lines.push_back(e);
        dfs.push_back(df);
    }
    std::set<int> nset;
    for (const auto &e : d.en) nset.insert(e.node);
    for (const auto &k : d.kk) nset.insert(k.first);
    for (int e : lines) {
        nset.insert(net.bus0[e]);
        nset.insert(net.bus1[e]);
    }
    for (const auto &e : d.en) {
        mt.en_node.push_back(e.node);
        mt.en_c.push_back(e.c);
        mt.en_div.push_back(e.div);
    }
    for (const auto &k : d.kk) {
        mt.kk_node.push_back(k.first);
        mt.kk_g.push_back(k.second);
    }
    for (size_t l = 0; l < lines.size(); l++) {
        mt.ln_line.push_back(lines[l]);
        mt.ln_df.push_back(dfs[l]);
    }
    for (int n : nset) {
        int ent = -1, kk = -1;
        for (size_t q = 0; q < d.en.size(); q++)
            if (d.en[q].node == n) ent = (int)q;
        for (size_t q = 0; q < d.kk.size(); q++)
            if (d.kk[q].first == n) kk = (int)q;
        mt.nd_node.push_back(n);
        mt.nd_ent.push_back(ent);
        mt.nd_kk.push_back(kk);
        for (size_t l = 0; l < lines.size(); l++) {
            const int e = lines[l];
            if (net.bus0[e] == n || net.bus1[e] == n) {
                mt.nl_pos.push_back((int)l);
                mt.nl_sgn.push_back(net.bus0[e] == n ? 1 : -1);
            }
        }
        mt.nd_l.push_back((int)mt.nl_pos.size());
    }
    mt.th_off.push_back((int)mt.th_node.size());
    mt.en_off.push_back((int)mt.en_node.size());
    mt.kk_off.push_back((int)mt.kk_node.size());
    mt.ln_off.push_back((int)mt.ln_line.size());
    mt.nd_off.push_back((int)mt.nd_node.size());
    mt.ntpl++;
}

// nodes within R hops of n (truncated breadth-first search over the neighbour lists)
std::vector<int> hops(const std::vector<std::vector<int>> &nbr, int n, int R) {
    std::vector<int> out{n};
    std::map<int, int> dist{{n, 0}};
    for (size_t head = 0; head < out.size(); head++) {
        const int i = out[head];
        if (dist[i] >= R) continue;
        for (int j : nbr[i]) {
            if (dist.count(j)) continue;
            dist[j] = dist[i] + 1;
            out.push_back(j);
        }
    }
    return out;
}

}  // namespace

bool build_move_tables(const esm_handle *h, MoveTablesHost &mt, std::string &err) {
    Net net;
    net.N = h->N;
    net.L = h->L;
    net.bus0 = h->bus0.data();
    net.bus1 = h->bus1.data();
    net.b = h->line_b.data();
    const int N = h->N, L = h->L;
    net.nodelines.assign(N, {});
    std::vector<std::vector<int>> nbr(N);
    for (int e = 0; e < L; e++) {
        const int a = net.bus0[e], b = net.bus1[e];
        net.nodelines[a].push_back(e);
        net.nodelines[b].push_back(e);
        if (a == b) continue;
        nbr[a].push_back(b);
        nbr[b].push_back(a);
    }
    // components (increasing node order inside each)
    net.comp.assign(N, -1);
    for (int s = 0; s < N; s++) {
        if (net.comp[s] >= 0) continue;
        const int c = (int)net.cnodes.size();
        std::vector<int> stack{s}, nodes;
        net.comp[s] = c;
        while (!stack.empty()) {
            const int i = stack.back();
            stack.pop_back();
            nodes.push_back(i);
            for (int j : nbr[i])
                if (net.comp[j] < 0) {
                    net.comp[j] = c;
                    stack.push_back(j);
                }
        }
        std::sort(nodes.begin(), nodes.end());
        net.cnodes.push_back(nodes);
    }
    // grounded Laplacian factor of every component (ground: its first node)
    std::vector<std::vector<double>> Gc(net.cnodes.size());
    std::vector<int> loc(N, 0);
    for (size_t c = 0; c < net.cnodes.size(); c++) {
        const auto &nd = net.cnodes[c];
        for (size_t q = 1; q < nd.size(); q++) loc[nd[q]] = (int)q;
        loc[nd[0]] = 0;
        const int k = (int)nd.size() - 1;
        if (k == 0) continue;
        std::vector<double> A((size_t)k * k, 0.0);
        for (int e = 0; e < L; e++) {
            const int a = net.bus0[e], b = net.bus1[e];
            if (net.comp[a] != (int)c) continue;
            const int ia = loc[a], ib = loc[b];
            const double w = net.b[e];
            if (ia > 0) A[(size_t)(ia - 1) * k + (ia - 1)] += w;
            if (ib > 0) A[(size_t)(ib - 1) * k + (ib - 1)] += w;
            if (ia > 0 && ib > 0) {
                A[(size_t)(ia - 1) * k + (ib - 1)] -= w;
                A[(size_t)(ib - 1) * k + (ia - 1)] -= w;
            }
        }
        if (!cholesky(A, k)) {
            err = "move tables: the grounded line Laplacian is not positive definite";
            return false;
        }
        Gc[c] = std::move(A);
    }
    // exact re-routing on the component of n: Δθ with L Δθ = e_n − e_m, grounded
    auto reroute = [&](int n, int m, std::vector<double> &x) {
        const int c = net.comp[n];
        const auto &nd = net.cnodes[c];
        const int k = (int)nd.size() - 1;
        std::vector<double> rhs(k, 0.0);
        if (loc[n] > 0) rhs[loc[n] - 1] += 1.0;
        if (loc[m] > 0) rhs[loc[m] - 1] -= 1.0;
        if (k > 0) chol_solve(Gc[c], k, rhs.data());
        x.assign(N, 0.0);
        for (int i : nd) x[i] = loc[i] > 0 ? rhs[loc[i] - 1] : 0.0;
    };

    mt = MoveTablesHost();
    mt.th_off.push_back(0);
    mt.en_off.push_back(0);
    mt.kk_off.push_back(0);
    mt.ln_off.push_back(0);
    mt.nd_off.push_back(0);
    mt.nd_l.push_back(0);
    // xloc / xlocc: partners within one hop, ball radius 0 (src/run.jl:23-24 defaults)
    mt.xp_off.assign(N + 1, 0);
    std::vector<double> x;
    for (int n = 0; n < N; n++) {
        std::vector<int> part = hops(nbr, n, 1);
        part.erase(std::remove(part.begin(), part.end(), n), part.end());
        std::sort(part.begin(), part.end());
        for (int m : part) {
            Dir d;
            d.en.push_back({n, 1.0, 1});
            d.en.push_back({m, -1.0, 1});
            std::vector<int> B = {std::min(n, m), std::max(n, m)};
            if ((int)B.size() >= (int)net.cnodes[net.comp[n]].size()) {
                reroute(n, m, x);
                for (int i : net.cnodes[net.comp[n]])
                    if (fabs(x[i]) >= 1e-15) d.th.push_back({i, x[i]});
            } else {
                const int nb = (int)B.size();
                std::vector<double> A((size_t)nb * nb, 0.0);
                auto pos = [&](int i) {
                    for (int q = 0; q < nb; q++)
                        if (B[q] == i) return q;
                    return -1;
                };
                for (int q = 0; q < nb; q++)
                    for (int e : net.nodelines[B[q]]) {
                        const int a = net.bus0[e], b = net.bus1[e];
                        const int o = a == B[q] ? b : a;
                        A[(size_t)q * nb + q] += net.b[e];
                        const int po = pos(o);
                        if (po >= 0) A[(size_t)q * nb + po] -= net.b[e];
                    }
                if (!cholesky(A, nb)) {
                    err = "move tables: a ball Laplacian is not positive definite";
                    return false;
                }
                std::vector<double> th(nb, 0.0);
                th[pos(n)] += 1.0;
                th[pos(m)] -= 1.0;
                chol_solve(A, nb, th.data());
                for (int q = 0; q < nb; q++)
                    if (fabs(th[q]) >= 1e-15) d.th.push_back({B[q], th[q]});
                // leak: requirement change −Σ b Δθ_i at the first layer outside B
                std::vector<std::pair<int, double>> lk;
                for (int q = 0; q < nb; q++)
                    for (int e : net.nodelines[B[q]]) {
                        const int a = net.bus0[e], b = net.bus1[e];
                        const int o = a == B[q] ? b : a;
                        if (pos(o) >= 0) continue;
                        size_t jl = 0;
                        while (jl < lk.size() && lk[jl].first != o) jl++;
                        if (jl == lk.size()) lk.push_back({o, 0.0});
                        lk[jl].second -= net.b[e] * th[q];
                    }
                std::sort(lk.begin(), lk.end());
                for (const auto &p : lk)
                    if (p.second != 0.0) d.kk.push_back(p);
            }
            emit(mt, net, d);
        }
        mt.xp_off[n + 1] = mt.ntpl;
    }
    // xpair: every ordered pair (n, m), n ≠ m in one component, global re-routing
    mt.xp0 = mt.ntpl;
    for (int n = 0; n < N; n++)
        for (int m = 0; m < N; m++) {
            Dir d;
            if (m != n && net.comp[m] == net.comp[n]) {
                d.en.push_back({n, 1.0, 1});
                d.en.push_back({m, -1.0, 1});
                reroute(n, m, x);
                for (int i : net.cnodes[net.comp[n]])
                    if (fabs(x[i]) >= 1e-15) d.th.push_back({i, x[i]});
            }
            emit(mt, net, d);
        }
    // cells of every partition level: the whole system, the countries, the nodes
    mt.cl0 = mt.ntpl;
    std::vector<std::vector<int>> labels;
    labels.push_back(std::vector<int>(N, 0));
    if (h->ncountry > 0) labels.push_back(h->country);
    std::vector<int> self(N);
    for (int n = 0; n < N; n++) self[n] = n;
    labels.push_back(self);
    mt.nlev = (int)labels.size();
    mt.lev_off.push_back(0);
    std::vector<double> qw;
    for (const auto &lab : labels) {
        std::vector<int> keys;   // cells in order of first appearance
        for (int n = 0; n < N; n++)
            if (std::find(keys.begin(), keys.end(), lab[n]) == keys.end()) keys.push_back(lab[n]);
        for (int key : keys) {
            Dir d;
            for (int n = 0; n < N; n++)
                if (lab[n] == key) d.en.push_back({n, 1.0, 0});
            emit(mt, net, d);
        }
        mt.lev_off.push_back(mt.lev_off.back() + (int)keys.size());
        qw.push_back((double)keys.size());
    }
    double acc = 0.0;
    for (double w : qw) {
        acc += w;
        mt.qcum.push_back(acc);
    }
    for (double &q : mt.qcum) q /= acc;
    // storage pairs (`sx_templates`, src/storage_cross.jl:29-53): per node its own slot, then per
    // neighbour m (increasing) the unit dipole n → m re-routed on the ball of radius one around the
    // pair (`xloc_directions` with rb = 1, src/local.jl:122-191; the exact global re-routing when
    // the ball holds the whole component), its leak into the first layer outside the ball
    {
        mt.sx_poff.assign(1, 0);
        for (auto *v : {&mt.sx_th_off, &mt.sx_kk_off, &mt.sx_ln_off, &mt.sx_nd_off, &mt.sx_nd_l, &mt.sx_ft_off})
            v->assign(1, 0);
        auto slot = [&](int n, int m, const std::vector<std::pair<int, double>> &th,
                        const std::vector<std::pair<int, double>> &lk) {
            std::vector<double> dth(N, 0.0);
            std::vector<char> lmark(L, 0), nmark(N, 0), fmark(N, 0);
            for (const auto &q : th) {
                mt.sx_th_node.push_back(q.first);
                mt.sx_th_val.push_back(q.second);
                dth[q.first] = q.second;
                fmark[q.first] = 1;
                for (int e : net.nodelines[q.first]) {
                    lmark[e] = 1;
                    fmark[net.bus0[e]] = fmark[net.bus1[e]] = 1;
                }
            }
            std::vector<int> ll;
            for (int e = 0; e < L; e++) {
                if (!lmark[e]) continue;
                const double ta = dth[net.bus0[e]], tb = dth[net.bus1[e]];
                if (ta == tb) continue;   // no flow change at any scale
                mt.sx_ln_line.push_back(e);
                mt.sx_ln_ta.push_back(ta);
                mt.sx_ln_tb.push_back(tb);
                ll.push_back(e);
                nmark[net.bus0[e]] = nmark[net.bus1[e]] = 1;
            }
            for (const auto &q : lk) {
                mt.sx_kk_node.push_back(q.first);
                mt.sx_kk_g.push_back(q.second);
                nmark[q.first] = fmark[q.first] = 1;
            }
            nmark[n] = nmark[m] = fmark[n] = fmark[m] = 1;
            for (int i = 0; i < N; i++) {
                if (!nmark[i]) continue;
                int kk = -1;
                for (size_t q = 0; q < lk.size(); q++)
                    if (lk[q].first == i) kk = (int)q;
                mt.sx_nd_node.push_back(i);
                mt.sx_nd_kk.push_back(kk);
                for (size_t l = 0; l < ll.size(); l++) {
                    const int e = ll[l];
                    if (net.bus0[e] == i || net.bus1[e] == i) {
                        mt.sx_nl_pos.push_back((int)l);
                        mt.sx_nl_sgn.push_back(net.bus0[e] == i ? 1 : -1);
                    }
                }
                mt.sx_nd_l.push_back((int)mt.sx_nl_pos.size());
            }
            for (int i = 0; i < N; i++)
                if (fmark[i]) mt.sx_ft_node.push_back(i);
            mt.sx_part.push_back(m);
            mt.sx_th_off.push_back((int)mt.sx_th_node.size());
            mt.sx_kk_off.push_back((int)mt.sx_kk_node.size());
            mt.sx_ln_off.push_back((int)mt.sx_ln_line.size());
            mt.sx_nd_off.push_back((int)mt.sx_nd_node.size());
            mt.sx_ft_off.push_back((int)mt.sx_ft_node.size());
        };
        for (int n = 0; n < N; n++) {
            slot(n, n, {}, {});
            std::vector<int> part = hops(nbr, n, 1);
            part.erase(std::remove(part.begin(), part.end(), n), part.end());
            std::sort(part.begin(), part.end());
            for (int m : part) {
                std::vector<int> B = hops(nbr, n, 1), Bm = hops(nbr, m, 1);
                B.insert(B.end(), Bm.begin(), Bm.end());
                std::sort(B.begin(), B.end());
                B.erase(std::unique(B.begin(), B.end()), B.end());
                std::vector<std::pair<int, double>> th, lk;
                if (B.size() >= net.cnodes[net.comp[n]].size()) {
                    reroute(n, m, x);
                    for (int i : net.cnodes[net.comp[n]])
                        if (fabs(x[i]) >= 1e-15) th.push_back({i, x[i]});
                } else {
                    const int nb = (int)B.size();
                    std::vector<int> pos(N, -1);
                    for (int q = 0; q < nb; q++) pos[B[q]] = q;
                    std::vector<double> A((size_t)nb * nb, 0.0);
                    for (int q = 0; q < nb; q++)
                        for (int e : net.nodelines[B[q]]) {
                            const int o = net.bus0[e] == B[q] ? net.bus1[e] : net.bus0[e];
                            A[(size_t)q * nb + q] += net.b[e];
                            if (pos[o] >= 0) A[(size_t)q * nb + pos[o]] -= net.b[e];
                        }
                    if (!cholesky(A, nb)) {
                        err = "move tables: a storage-pair ball Laplacian is not positive definite";
                        return false;
                    }
                    std::vector<double> t(nb, 0.0);
                    t[pos[n]] += 1.0;
                    t[pos[m]] -= 1.0;
                    chol_solve(A, nb, t.data());
                    for (int q = 0; q < nb; q++)
                        if (fabs(t[q]) >= 1e-15) th.push_back({B[q], t[q]});
                    for (int q = 0; q < nb; q++)
                        for (int e : net.nodelines[B[q]]) {
                            const int o = net.bus0[e] == B[q] ? net.bus1[e] : net.bus0[e];
                            if (pos[o] >= 0) continue;
                            size_t jl = 0;
                            while (jl < lk.size() && lk[jl].first != o) jl++;
                            if (jl == lk.size()) lk.push_back({o, 0.0});
                            lk[jl].second -= net.b[e] * t[q];
                        }
                    std::sort(lk.begin(), lk.end(),
                              [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
                }
                slot(n, m, th, lk);
            }
            mt.sx_poff.push_back((int)mt.sx_part.size());
        }
    }
    // storage multiscale moves (`storage_ms_sweep!`, src/flowmoves.jl:266-277): the units, and
    // the groups of one carrier in one country (cellof = E.parts[2]; single nodes without countries)
    for (int s = 0; s < ESM_NSTOR; s++)
        for (int n = 0; n < N; n++) mt.sms_units.push_back(s * N + n);
    std::vector<int> cellof(N);
    for (int n = 0; n < N; n++) cellof[n] = h->ncountry > 0 ? h->country[n] : n;
    std::vector<int> cells(cellof);
    std::sort(cells.begin(), cells.end());
    cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
    mt.grp_off.push_back(0);
    for (int c : cells)
        for (int s = 0; s < ESM_NSTOR; s++) {
            int cnt = 0;
            for (int n = 0; n < N; n++)
                if (cellof[n] == c) {
                    mt.grp_node.push_back(n);
                    cnt++;
                }
            if (cnt > SMS_NU) {
                err = "move tables: a country holds more than " + std::to_string(SMS_NU) +
                      " nodes (the storage group limit: one lane stages each unit)";
                return false;
            }
            mt.grp_si.push_back(s);
            mt.grp_off.push_back((int)mt.grp_node.size());
        }
    return true;
}

}  // namespace esm
