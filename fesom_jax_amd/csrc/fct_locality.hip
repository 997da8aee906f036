// Edge patches of the adf_h patch kernel, built on the host with the plan.
//
// The caller's node and edge numbering is kept for every array (results are written
// where the caller expects them, and every per-node sum still runs over the node's
// (edge, side) list in increasing 2*edge + side, the reference's scatter order). Only
// the adf_h patch kernel visits edges in another order: nodes are grouped into compact
// balls grown breadth-first, each ball takes the edges whose first node it holds, and
// the balls are grouped the same way twice more (up to 16 per group, then up to 16
// groups per super-group), so that the edges one workgroup stages together share their
// node columns whatever the numbering. The node passes and the vertical-flux kernel keep
// the id order: on the meshes measured here the patch order made them slower.

#include <algorithm>
#include <vector>

#include "fct_common.h"

namespace fct {

namespace {

// Breadth-first growth of groups over a graph of `n` items (CSR adjacency `off`/`adj`):
// a group takes items in breadth-first order from its seed while their weights fit
// `cap`; the next seed is the oldest unassigned item next to an earlier group. Returns
// the groups as consecutive runs of `order`, delimited by `start`.
void grow_groups(int n, const std::vector<int>& off, const std::vector<int>& adj,
                 const std::vector<int>& weight, int cap, std::vector<int>& order,
                 std::vector<int>& start) {
    std::vector<char> taken(n, 0);
    std::vector<int> front, queue;
    order.clear();
    start.clear();
    order.reserve(n);
    size_t fhead = 0;
    int scan = 0;
    while ((int)order.size() < n) {
        int seed = -1;
        while (fhead < front.size() && seed < 0) {
            const int c = front[fhead++];
            if (!taken[c]) seed = c;
        }
        while (seed < 0) {
            if (!taken[scan]) seed = scan;
            ++scan;
        }
        start.push_back((int)order.size());
        queue.clear();
        queue.push_back(seed);
        taken[seed] = 1;
        order.push_back(seed);
        int total = weight[seed];
        for (size_t qi = 0; qi < queue.size(); ++qi) {
            const int u = queue[qi];
            for (int k = off[u]; k < off[u + 1]; ++k) {
                const int v = adj[k];
                if (taken[v] || total + weight[v] > cap) continue;
                taken[v] = 1;
                total += weight[v];
                order.push_back(v);
                queue.push_back(v);
            }
        }
        for (int u : queue)
            for (int k = off[u]; k < off[u + 1]; ++k)
                if (!taken[adj[k]]) front.push_back(adj[k]);
    }
    start.push_back(n);
}

}  // namespace

int build_locality(int N, int D, const int2* edges, const int* eoff, const int* ecsr,
                   const int4* udv, const unsigned char* has_updn, const int2* ekr,
                   const int2* utri, LocalityHost& out) {
    // edges owned by each node: those whose first node it is
    std::vector<int> own_off(N + 1, 0), own(D);
    for (int e = 0; e < D; ++e) own_off[edges[e].x + 1]++;
    for (int n = 0; n < N; ++n) own_off[n + 1] += own_off[n];
    {
        std::vector<int> pos(own_off.begin(), own_off.end() - 1);
        for (int e = 0; e < D; ++e) own[pos[edges[e].x]++] = e;
    }

    // 1. node patches: breadth-first balls of nodes owning at most FCT_EP edges; then
    //    two coarser levels, each grouping up to 16 adjacent groups of the level below,
    //    so that the patch order keeps neighbours close at every scale (a ball's
    //    neighbours in its group, a group's in its super-group)
    std::vector<int> noff(N + 1), nadj(eoff[N]), nweight(N);
    for (int n = 0; n <= N; ++n) noff[n] = eoff[n];
    for (int n = 0; n < N; ++n) {
        nweight[n] = own_off[n + 1] - own_off[n];
        for (int k = eoff[n]; k < eoff[n + 1]; ++k) {
            const int2 nd = edges[ecsr[k] >> 1];
            nadj[k] = (ecsr[k] & 1) ? nd.x : nd.y;
        }
    }
    std::vector<int> order, ball_start;
    grow_groups(N, noff, nadj, nweight, FCT_EP, order, ball_start);
    // groups[g] = the node-order runs (as ball indices) of each group, finest first
    std::vector<std::vector<int>> members(ball_start.size() - 1);
    for (size_t b = 0; b + 1 < ball_start.size(); ++b) members[b] = {(int)b};
    std::vector<int> owner(N);
    for (size_t b = 0; b + 1 < ball_start.size(); ++b)
        for (int i = ball_start[b]; i < ball_start[b + 1]; ++i) owner[order[i]] = (int)b;
    for (int level = 0; level < 2; ++level) {
        const int G = (int)members.size();
        std::vector<std::vector<int>> nb(G);
        for (int e = 0; e < D; ++e) {
            const int a = owner[edges[e].x], b = owner[edges[e].y];
            if (a != b) { nb[a].push_back(b); nb[b].push_back(a); }
        }
        std::vector<int> goff(G + 1, 0), gadj, gw(G, 1), gorder, gstart;
        for (int g = 0; g < G; ++g) {
            std::sort(nb[g].begin(), nb[g].end());
            nb[g].erase(std::unique(nb[g].begin(), nb[g].end()), nb[g].end());
            goff[g + 1] = goff[g] + (int)nb[g].size();
            gadj.insert(gadj.end(), nb[g].begin(), nb[g].end());
        }
        grow_groups(G, goff, gadj, gw, 16, gorder, gstart);
        std::vector<std::vector<int>> next(gstart.size() - 1);
        std::vector<int> gowner(G);
        for (size_t c = 0; c + 1 < gstart.size(); ++c)
            for (int i = gstart[c]; i < gstart[c + 1]; ++i) {
                const int g = gorder[i];
                gowner[g] = (int)c;
                next[c].insert(next[c].end(), members[g].begin(), members[g].end());
            }
        for (int n = 0; n < N; ++n) owner[n] = gowner[owner[n]];
        members.swap(next);
    }
    std::vector<int> balls;   // balls in the final order
    for (const auto& g : members) balls.insert(balls.end(), g.begin(), g.end());
    out.node_order.clear();
    out.node_order.reserve(N);
    for (int b : balls)
        out.node_order.insert(out.node_order.end(), order.begin() + ball_start[b],
                              order.begin() + ball_start[b + 1]);

    // 2. edge patches: the owned edges of each node patch, split where the node list
    //    (edge ends and up/down triangle vertices) would pass FCT_PU
    out.ptab.clear(); out.pnodes.clear(); out.elist.clear(); out.eslot.clear();
    out.elist.reserve(D); out.eslot.reserve(D);
    std::vector<int> slot(N, -1), cur_nodes, cur_edges;
    auto flush = [&]() {
        if (cur_edges.empty()) return;
        // wet levels of the patch, and the first level at which one of its edges reads
        // the node-averaged gradients (outside its up/down triangle range)
        int klo = 1 << 30, khi = 0, kg = 1 << 30;
        for (int e : cur_edges) {
            if (ekr[e].x >= ekr[e].y) continue;
            klo = std::min(klo, ekr[e].x);
            khi = std::max(khi, ekr[e].y);
            const int2 u = utri[e];
            const bool full = u.x <= ekr[e].x && u.y >= ekr[e].y && u.x < u.y;
            if (!full) kg = std::min(kg, (u.x < u.y && u.x <= ekr[e].x) ? u.y : ekr[e].x);
        }
        if (klo >= khi) klo = khi = 0;
        kg = std::min(kg, khi);
        out.ptab.push_back(make_int4((int)out.elist.size(), (int)out.pnodes.size(),
                                     (int)cur_edges.size() | ((int)cur_nodes.size() << 16),
                                     klo | (khi << 8) | (kg << 16)));
        for (int e : cur_edges) {
            const int2 nd = edges[e];
            const int4 v = udv[2 * (size_t)e];
            const int cc = udv[2 * (size_t)e + 1].x;
            const bool t = has_updn[e] != 0;
            const int s1 = slot[nd.x], s2 = slot[nd.y];
            const int ua = t ? slot[v.x] : s1, ub = t ? slot[v.y] : s1;
            const int da = t ? slot[v.z] : s2, db = t ? slot[v.w] : s2;
            out.elist.push_back(e);
            out.eslot.push_back(make_int2(s1 | (s2 << 8) | (ua << 16) | (ub << 24),
                                          da | (db << 8) | ((cc & 3) << 16) | (((cc >> 2) & 3) << 18)));
        }
        for (int n : cur_nodes) { out.pnodes.push_back(n); slot[n] = -1; }
        cur_nodes.clear();
        cur_edges.clear();
    };
    for (int b : balls) {
        for (int i = ball_start[b]; i < ball_start[b + 1]; ++i) {
            const int n = order[i];
            for (int k = own_off[n]; k < own_off[n + 1]; ++k) {
                const int e = own[k];
                const int2 nd = edges[e];
                const int4 v = udv[2 * (size_t)e];
                int need[6] = {nd.x, nd.y, v.x, v.y, v.z, v.w};
                const int nneed = has_updn[e] ? 6 : 2;
                int fresh = 0;
                for (int q = 0; q < nneed; ++q) {
                    bool seen = slot[need[q]] >= 0;
                    for (int r = 0; r < q && !seen; ++r) seen = need[r] == need[q];
                    fresh += !seen;
                }
                if ((int)cur_nodes.size() + fresh > FCT_PU || (int)cur_edges.size() == FCT_EP)
                    flush();
                for (int q = 0; q < nneed; ++q)
                    if (slot[need[q]] < 0) {
                        slot[need[q]] = (int)cur_nodes.size();
                        cur_nodes.push_back(need[q]);
                    }
                cur_edges.push_back(e);
            }
        }
        flush();
    }
    if ((int)out.elist.size() != D)
        return set_error(FCT_E_INVALID, "locality order lost edges (%zu of %d)", out.elist.size(), D);
    return FCT_OK;
}

}  // namespace fct
