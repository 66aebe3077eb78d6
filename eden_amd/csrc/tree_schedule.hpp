// tree_schedule.hpp -- host-side planning for the tree-cell kernel: breadth-first relabelling
// and round schedules of the Hines solve, and the launch shape per cell size.
#pragma once
#include <algorithm>
#include <cmath>
#include <queue>
#include <stdint.h>
#include <utility>
#include <vector>

#include "../../include/edenhip.h"

namespace edenhip {

// Rate value in double, used only for the setup value of gate states (alpha/(alpha+beta) at
// V0, eden/Eden.cpp:5130-5148, 5313-5321), which the initialisation steps overwrite.
inline double edenhip_rate_host(const edenhip_rate &r, double V) {
    switch (r.form) {
    case EDENHIP_RATE_EXP: return r.rate * std::exp((V - r.midpoint) / r.scale);
    case EDENHIP_RATE_EXPLINEAR: {
        double x = (V - r.midpoint) / r.scale;
        return x == 0 ? r.rate : r.rate * x / (1 - std::exp(-x));
    }
    case EDENHIP_RATE_SIGMOID: return r.rate / (1 + std::exp((r.midpoint - V) / r.scale));
    default: return r.rate;
    }
}

constexpr int32_t MAX_TREE_N = 2047;   // 11-bit positions in task words, plus the dummy slot

// Schedule of one tree for a workgroup of `tpb` threads.
//
// Positions: nodes are relabelled breadth-first from the root, visiting each node's children
// in ascending index order, so a node's children occupy consecutive positions in that order.
//
// Forward elimination runs over the nodes that have children, grouped by height (1 = only
// leaf children); each such node folds its children into itself in ascending child index,
// the order in which the reference's depth-first post-order elimination applies them
// (adjacency lists [parent, children ascending], eden/Eden.cpp:4070-4078, 4416-4486), so the
// diagonal and right-hand side of every node see the same operations in the same order.
// Back-substitution runs over non-root nodes grouped by depth.
//
// Each group is split into rounds of at most `tpb` tasks, one 32-bit word per task:
//   forward:  pos | first_child_pos << 11 | n_children << 22
//   backward: pos | parent_pos << 11
// The schedule of a cell is one block of 32-bit words: the round headers (offset of the
// round's first word in the block | task count << 16), forward rounds followed by one empty
// round, backward rounds followed by one empty round, then the task words. Threads at or
// beyond a round's task count take an idle word that points every access at the dummy slot
// n (forward: node n with one child n; backward: node n with parent n), so the rounds run
// as straight-line code; the empty rounds are read as the prefetch past a sweep's end.
struct TreePlan {
    std::vector<int32_t> pos;          // [n] BFS position of each compartment
    std::vector<int32_t> parent_pos;   // [n] by position
    int32_t rounds_fwd = 0, rounds_bwd = 0;
    std::vector<uint32_t> block;       // headers (rounds_fwd + rounds_bwd + 2), then task words
};

inline uint32_t idle_fwd_word(int32_t n) { return (uint32_t)n | ((uint32_t)n << 11) | (1u << 22); }
inline uint32_t idle_bwd_word(int32_t n) { return (uint32_t)n | ((uint32_t)n << 11); }

inline TreePlan plan_tree(int32_t n, const int32_t *parent, int32_t tpb) {
    TreePlan P;
    std::vector<std::vector<int32_t>> ch(n);
    for (int32_t i = 1; i < n; i++) ch[parent[i]].push_back(i);   // ascending
    P.pos.assign(n, -1);
    std::vector<int32_t> order;
    order.reserve(n);
    order.push_back(0);
    for (size_t h = 0; h < order.size(); h++)
        for (int32_t c : ch[order[h]]) order.push_back(c);
    for (int32_t p = 0; p < n; p++) P.pos[order[p]] = p;
    P.parent_pos.assign(n, -1);
    for (int32_t i = 1; i < n; i++) P.parent_pos[P.pos[i]] = P.pos[parent[i]];
    std::vector<int32_t> height(n, 0), depth(n, 0);
    for (int32_t i = n - 1; i >= 1; i--) height[parent[i]] = std::max(height[parent[i]], height[i] + 1);
    for (int32_t i = 1; i < n; i++) depth[i] = depth[parent[i]] + 1;
    const int32_t H = height[0];
    int32_t Dm = 0;
    for (int32_t i = 0; i < n; i++) Dm = std::max(Dm, depth[i]);
    std::vector<uint32_t> words;
    std::vector<std::pair<uint32_t, uint32_t>> rounds;   // (first word, count)
    auto emit = [&](const std::vector<uint32_t> &tasks, int32_t &nr) {
        for (size_t k = 0; k < tasks.size(); k += (size_t)tpb) {
            const size_t cnt = std::min(tasks.size() - k, (size_t)tpb);
            rounds.emplace_back((uint32_t)words.size(), (uint32_t)cnt);
            words.insert(words.end(), tasks.begin() + (long)k, tasks.begin() + (long)(k + cnt));
            nr++;
        }
    };
    for (int32_t h = 1; h <= H; h++) {
        std::vector<uint32_t> tasks;
        for (int32_t i = 0; i < n; i++)
            if (height[i] == h) {
                const uint32_t p = (uint32_t)P.pos[i];
                const uint32_t fc = (uint32_t)P.pos[ch[i][0]];
                tasks.push_back(p | (fc << 11) | ((uint32_t)ch[i].size() << 22));
            }
        emit(tasks, P.rounds_fwd);
    }
    rounds.emplace_back((uint32_t)words.size(), 0u);
    for (int32_t d = 1; d <= Dm; d++) {
        std::vector<uint32_t> tasks;
        for (int32_t p = 0; p < n; p++) {
            const int32_t i = order[p];
            if (depth[i] == d) tasks.push_back((uint32_t)p | ((uint32_t)P.parent_pos[p] << 11));
        }
        emit(tasks, P.rounds_bwd);
    }
    rounds.emplace_back((uint32_t)words.size(), 0u);
    const uint32_t nh = (uint32_t)rounds.size();
    for (auto &r : rounds) P.block.push_back((nh + r.first) | (r.second << 16));
    P.block.insert(P.block.end(), words.begin(), words.end());
    return P;
}

// Wave-partitioned schedule of one tree for a workgroup of `waves` cable waves (tree kernel).
//
// Positions as in plan_tree (breadth-first, children consecutive in ascending index order),
// and the same per-node operations: a forward task folds all children of one node in
// ascending child index, a backward task solves one non-root node from its parent.
//
// The tree is cut into a junction J (the root and, for several waves, a few nodes below it:
// J is closed towards the root) and the subtrees hanging from J, which are dealt to the
// waves. A wave runs the forward tasks of its subtrees by height, then, after one workgroup
// barrier, wave 0 runs the forward tasks of J (by height within J), solves the root and runs
// the backward tasks of J's non-root nodes (by depth); after a second barrier every wave runs
// the backward tasks of its subtrees by depth. Within a wave the rounds follow one another
// with no barrier: a wave's LDS accesses are performed in program order, so a round reads
// what the previous rounds of the same wave wrote. With one wave, J is empty and the root's
// forward task is the last of wave 0's forward rounds.
//
// Block layout (32-bit words):
//   directory  [2 w] = wave w's forward list, [2 w + 1] = its backward list (w < waves),
//              [2 waves] = J's forward list, [2 waves + 1] = J's backward list;
//              a list is (index of its first round header | round count << 16)
//   headers    one per round: (offset of its first task word | task count << 16), task
//              count 1..64, then five empty sentinel headers
//   words      forward: pos | first_child_pos << 11 | n_children << 22
//              backward: pos | parent_pos << 11
// The five rounds after a list's last round are always headers (of the next lists or the
// sentinels), so a wave may read up to five headers past its list.
struct WavePlan {
    std::vector<int32_t> pos;          // [n] BFS position of each compartment
    int32_t n_junction = 0;            // nodes in J
    int32_t longest = 0;               // rounds on the longest path (max wave + J, both sweeps)
    std::vector<uint32_t> block;
};

inline WavePlan plan_tree_waves(int32_t n, const int32_t *parent, int32_t waves) {
    WavePlan P;
    std::vector<std::vector<int32_t>> ch(n);
    for (int32_t i = 1; i < n; i++) ch[parent[i]].push_back(i);   // ascending
    std::vector<int32_t> order;
    order.reserve(n);
    order.push_back(0);
    for (size_t h = 0; h < order.size(); h++)
        for (int32_t c : ch[order[h]]) order.push_back(c);
    P.pos.assign(n, -1);
    for (int32_t p = 0; p < n; p++) P.pos[order[p]] = p;
    std::vector<int32_t> height(n, 0), depth(n, 0), size(n, 1);
    for (int32_t i = n - 1; i >= 1; i--) {
        height[parent[i]] = std::max(height[parent[i]], height[i] + 1);
        size[parent[i]] += size[i];
    }
    for (int32_t i = 1; i < n; i++) depth[i] = depth[parent[i]] + 1;
    constexpr int32_t LANES = 64;

    // rounds of a group of tasks keyed by level: one round per 64 tasks of a level
    auto count_rounds = [&](const std::vector<int32_t> &per_level) {
        int32_t r = 0;
        for (int32_t c : per_level) r += (c + LANES - 1) / LANES;
        return r;
    };
    // owner[i]: wave of node i, or -1 for J
    struct Cut {
        std::vector<int32_t> owner;
        int32_t cost = 0;
    };
    auto evaluate = [&](std::vector<int32_t> &owner) {
        const int32_t H = height[0] + 1, Dm = n;
        std::vector<std::vector<int32_t>> fw(waves, std::vector<int32_t>(H, 0)), bw(waves);
        int32_t fmax = 0, bmax = 0;
        for (int32_t w = 0; w < waves; w++) bw[w].assign((size_t)std::min(Dm, n), 0);
        for (int32_t i = 0; i < n; i++) {
            if (owner[i] < 0) continue;
            if (!ch[i].empty()) fw[owner[i]][height[i]]++;
            if (i > 0) bw[owner[i]][depth[i]]++;
        }
        for (int32_t w = 0; w < waves; w++) {
            fmax = std::max(fmax, count_rounds(fw[w]));
            bmax = std::max(bmax, count_rounds(bw[w]));
        }
        // J: forward levels by height within J, backward by depth
        std::vector<int32_t> jh(n, 0), jf(n + 1, 0), jb(n + 1, 0);
        for (int32_t k = n - 1; k >= 0; k--) {
            const int32_t i = order[k];
            if (owner[i] >= 0) continue;
            int32_t m = 0;
            for (int32_t c : ch[i])
                if (owner[c] < 0) m = std::max(m, jh[c]);
            jh[i] = m + 1;
            if (!ch[i].empty()) jf[jh[i]]++;
            if (i > 0) jb[depth[i]]++;
        }
        // the barriers cost about one round each
        return fmax + bmax + count_rounds(jf) + count_rounds(jb) + (waves > 1 ? 2 : 0);
    };
    Cut best;
    if (waves <= 1) {
        best.owner.assign(n, 0);
        best.cost = evaluate(best.owner);
    } else {
        const double share = (double)n / waves;
        for (double f : {0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 4.0, 1e9}) {
            const int32_t T = (int32_t)std::max(1.0, f * share);
            std::vector<int32_t> owner(n, 0);
            owner[0] = -1;
            // split the largest subtree while it exceeds T (a heap by size, ties by index)
            std::priority_queue<std::pair<int32_t, int32_t>> heap;
            for (int32_t c : ch[0]) heap.push({size[c], -c});
            std::vector<int32_t> cand;
            int32_t nj = 1;
            while (!heap.empty()) {
                const int32_t c = -heap.top().second;
                if (size[c] <= T || ch[c].empty() || nj >= 256) break;
                heap.pop();
                owner[c] = -1;
                nj++;
                for (int32_t x : ch[c]) heap.push({size[x], -x});
            }
            for (; !heap.empty(); heap.pop()) cand.push_back(-heap.top().second);
            // deal the subtrees, largest first, to the least loaded wave
            std::stable_sort(cand.begin(), cand.end(), [&](int32_t a, int32_t b) { return size[a] > size[b]; });
            std::vector<int64_t> load(waves, 0);
            for (int32_t c : cand) {
                const int32_t w = (int32_t)(std::min_element(load.begin(), load.end()) - load.begin());
                load[w] += size[c];
                std::vector<int32_t> st{c};
                while (!st.empty()) {
                    const int32_t v = st.back();
                    st.pop_back();
                    owner[v] = w;
                    for (int32_t x : ch[v]) st.push_back(x);
                }
            }
            const int32_t cost = evaluate(owner);
            if (best.owner.empty() || cost < best.cost) {
                best.owner = owner;
                best.cost = cost;
            }
        }
    }
    P.longest = best.cost;
    const std::vector<int32_t> &owner = best.owner;
    for (int32_t i = 0; i < n; i++) P.n_junction += owner[i] < 0 ? 1 : 0;

    // task lists: per wave, forward by height and backward by depth; then J
    std::vector<uint32_t> words;
    std::vector<std::pair<uint32_t, uint32_t>> rounds;   // (first word, count)
    std::vector<std::pair<uint32_t, uint32_t>> lists;    // (first round, round count)
    auto emit_list = [&](std::vector<std::vector<uint32_t>> &levels) {
        const uint32_t r0 = (uint32_t)rounds.size();
        for (auto &tasks : levels)
            for (size_t k = 0; k < tasks.size(); k += LANES) {
                const size_t cnt = std::min(tasks.size() - k, (size_t)LANES);
                rounds.emplace_back((uint32_t)words.size(), (uint32_t)cnt);
                words.insert(words.end(), tasks.begin() + (long)k, tasks.begin() + (long)(k + cnt));
            }
        lists.emplace_back(r0, (uint32_t)rounds.size() - r0);
    };
    auto fwd_word = [&](int32_t i) {
        return (uint32_t)P.pos[i] | ((uint32_t)P.pos[ch[i][0]] << 11) | ((uint32_t)ch[i].size() << 22);
    };
    auto bwd_word = [&](int32_t i) { return (uint32_t)P.pos[i] | ((uint32_t)P.pos[parent[i]] << 11); };
    const int32_t H = height[0];
    int32_t Dm = 0;
    for (int32_t i = 0; i < n; i++) Dm = std::max(Dm, depth[i]);
    for (int32_t w = 0; w < waves; w++) {
        std::vector<std::vector<uint32_t>> f(H + 1), b(Dm + 1);
        for (int32_t p = 0; p < n; p++) {   // position order within a level
            const int32_t i = order[p];
            if (owner[i] != w) continue;
            if (!ch[i].empty()) f[height[i]].push_back(fwd_word(i));
            if (i > 0) b[depth[i]].push_back(bwd_word(i));
        }
        emit_list(f);
        emit_list(b);
    }
    {
        std::vector<int32_t> jh(n, 0);
        for (int32_t k = n - 1; k >= 0; k--) {
            const int32_t i = order[k];
            if (owner[i] >= 0) continue;
            int32_t m = 0;
            for (int32_t c : ch[i])
                if (owner[c] < 0) m = std::max(m, jh[c]);
            jh[i] = m + 1;
        }
        std::vector<std::vector<uint32_t>> f(H + 2), b(Dm + 1);
        for (int32_t p = 0; p < n; p++) {
            const int32_t i = order[p];
            if (owner[i] >= 0) continue;
            if (!ch[i].empty()) f[jh[i]].push_back(fwd_word(i));
            if (i > 0) b[depth[i]].push_back(bwd_word(i));
        }
        emit_list(f);
        emit_list(b);
    }
    for (int k = 0; k < 5; k++) rounds.emplace_back((uint32_t)words.size(), 0u);   // sentinels
    const uint32_t nd = 2 * (uint32_t)waves + 2, nh = (uint32_t)rounds.size();
    for (auto &l : lists) P.block.push_back((nd + l.first) | (l.second << 16));
    for (auto &r : rounds) P.block.push_back((nd + nh + r.first) | (r.second << 16));
    P.block.insert(P.block.end(), words.begin(), words.end());
    return P;
}

// Workgroup size for a cell of n compartments: each thread holds up to 4 compartments.
inline int32_t choose_tpb(int32_t n, int32_t forced) {
    if (forced == 64 || forced == 128 || forced == 256 || forced == 512)
        if (n <= forced * 4) return forced;
    if (n <= 256) return 64;
    if (n <= 512) return 128;
    if (n <= 1024) return 256;
    return 512;
}

}  // namespace edenhip
