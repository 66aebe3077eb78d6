// tree_schedule.hpp -- host-side planning for the tree-cell kernel: breadth-first relabelling
// and round schedules of the Hines solve, and the launch shape per cell size.
#pragma once
#include <algorithm>
#include <cmath>
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
