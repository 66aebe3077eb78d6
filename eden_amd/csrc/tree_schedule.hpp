// tree_schedule.hpp -- host-side planning for the tree-cell kernel: breadth-first relabelling
// and round schedules of the Hines solve, and the launch shape per cell size.
#pragma once
#include <algorithm>
#include <array>
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
    bool ok = true;                    // false: a node has more than 1,023 children (10-bit count)
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
    for (int32_t i = 0; i < n; i++)
        if (ch[i].size() > 1023) P.ok = false;
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
// Positions as in plan_tree (breadth-first, children consecutive in ascending index order).
// A node with children is eliminated by fold tasks, each folding at most two consecutive
// children into the node's diagonal and right-hand side: children 2k and 2k + 1 in task k,
// the tasks of a node in ascending k. The node therefore sees the same operations in the
// same order as the reference's serial elimination, which applies its children one after
// the other in ascending index. The root's fold is not scheduled here: the kernel performs
// it (all root children in order) after every other fold. A backward task solves one
// non-root node from its solved parent, or, when it shares a round with its parent's task,
// solves the parent again from the solved grandparent and then itself: the same operations
// on the same values, so two depths take one round.
//
// The tree is cut into a junction J (the root and, for several waves, a few nodes below it:
// J is closed towards the root) and the subtrees hanging from J, which are dealt to the
// waves. A wave runs the fold tasks of its subtrees, then, after one workgroup barrier, wave
// 0 runs the fold tasks of J's non-root nodes, folds the root, solves it and runs the
// backward tasks of J's non-root nodes (by depth); after a second barrier every wave runs the
// backward tasks of its subtrees by depth. A fold task is placed in the round after the last
// of its inputs (the node's previous fold, the two children's last folds). Within a wave the
// rounds follow one another with no barrier: a wave's LDS accesses are performed in program
// order, so a round reads what the previous rounds of the same wave wrote. With one wave, J
// is empty (the root is folded after wave 0's rounds).
//
// Block layout (32-bit words):
//   directory  [2 w] = wave w's fold list, [2 w + 1] = its backward list (w < waves),
//              [2 waves] = J's fold list, [2 waves + 1] = J's backward list;
//              a list is (index of its first round header | round count << 16)
//   headers    one per round: (index of its first task word | task count << 16), task
//              count 1..64, then five empty sentinel headers; every round starts at an
//              index that is a multiple of the fold task's word count
//   words      byte offsets of node records, o(x) = WAVE_NODE_BYTES * position of x (the
//              dummy slot: o(n)). Fold task, two words: (o(pos) | o(first_child) << 16),
//              o(second child) (o(n) for a task with one child); with merge, four words:
//              (o(pos) | o(c0) << 16), (o(c1) | o(g00) << 16), (o(g01) | o(g10) << 16),
//              o(g11), where g00, g01 are the children of c0 when the task folds them into c0
//              again (merged child), else o(n), and g10, g11 those of c1. Backward task, two
//              words: (o(pos) | o(solved) << 16), o(mid), with solved the parent and mid =
//              o(n), or solved the grandparent and mid the parent
//   padding    zero words (wave_plan_pad), so that a wave may read 64 tasks from any round
// The five rounds after a list's last round are always headers (of the next lists or the
// sentinels), so a wave may read up to five headers past its list.
struct WavePlan {
    std::vector<int32_t> pos;          // [n] BFS position of each compartment
    int32_t n_junction = 0;            // nodes in J (the root included)
    int32_t hines_waves = 1;           // waves with non-empty lists
    int32_t longest = 0;               // rounds on the longest path (waves + J, both sweeps)
    bool merged = false;               // four-word fold tasks (merge)
    std::vector<uint32_t> block;
};

// bytes per node record of the tree kernel's Hines arrays: task words hold byte offsets
// WAVE_NODE_BYTES * position (16-bit fields: at most 16 * 2047)
constexpr uint32_t WAVE_NODE_BYTES = 16;

// zero words at the end of a block: 64 lanes reading one task each from the last round
inline int32_t wave_plan_pad(bool merge) { return merge ? 256 : 128; }

// first_bias, last_bias: initial loads of wave 0 and of the last wave, in shares of n / waves,
// for waves with other work in the same phase (the tree kernel's root compartment update on
// wave 0, gate updates on the last wave).
inline WavePlan plan_tree_waves(int32_t n, const int32_t *parent, int32_t waves, double first_bias = 0.0,
                                double last_bias = 0.0, bool merge = false) {
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
    std::vector<int32_t> depth(n, 0), size(n, 1);
    for (int32_t i = n - 1; i >= 1; i--) size[parent[i]] += size[i];
    for (int32_t i = 1; i < n; i++) depth[i] = depth[parent[i]] + 1;
    constexpr int32_t LANES = 64;

    // Fold tasks of the nodes of one list (owner == o), children before parents: task k of
    // node i folds children 2k and 2k + 1 and is placed at level 1 + max(level of task k - 1,
    // last level of those children); nodes outside the list (done in an earlier phase) and
    // leaves count as level 0. With merge, a child whose only fold task merged no child of its
    // own may instead share the task's level: the task then folds that child's children into
    // the child's value again in its own lane (the same operations on the same values as the
    // child's own task, which reads and writes in the same round) and folds the result, so
    // two tree levels take one round. emit(level, node, k, merged child 0, merged child 1).
    auto fold_tasks = [&](const std::vector<int32_t> &owner, int32_t o, auto &&emit) {
        std::vector<int32_t> last(n, 0);
        std::vector<char> single(n, 0);
        for (int32_t k = n - 1; k >= 0; k--) {   // reverse BFS: children before parents
            const int32_t i = order[k];
            if (owner[i] != o || i == 0 || ch[i].empty()) continue;
            int32_t lv = 0;
            bool merged_any = false;
            for (size_t c = 0; c < ch[i].size(); c += 2) {
                const size_t e = std::min(c + 2, ch[i].size());
                int32_t need = lv + 1;
                for (size_t d = c; d < e; d++) {
                    const int32_t x = ch[i][d];
                    if (owner[x] == o) need = std::max(need, merge && single[x] ? last[x] : last[x] + 1);
                }
                bool m[2] = {false, false};
                for (size_t d = c; d < e; d++) {
                    const int32_t x = ch[i][d];
                    m[d - c] = owner[x] == o && merge && single[x] && last[x] == need;
                    merged_any = merged_any || m[d - c];
                }
                emit(need, i, (int32_t)(c / 2), m[0], m[1]);
                lv = need;
            }
            last[i] = lv;
            single[i] = ch[i].size() <= 2 && !merged_any;
        }
    };
    auto fold_levels = [&](const std::vector<int32_t> &owner, int32_t o, std::vector<int32_t> &count) {
        fold_tasks(owner, o, [&](int32_t lv, int32_t, int32_t, bool, bool) {
            if ((int32_t)count.size() <= lv) count.resize(lv + 1, 0);
            count[lv]++;
        });
    };
    auto count_rounds = [&](const std::vector<int32_t> &per_level) {
        int32_t r = 0;
        for (int32_t c : per_level) r += (c + LANES - 1) / LANES;
        return r;
    };
    // backward rounds: two consecutive non-empty depths with at most 64 tasks share one
    auto bwd_rounds = [&](const std::vector<int32_t> &per_depth) {
        int32_t r = 0;
        for (size_t d = 0; d < per_depth.size(); d++) {
            if (per_depth[d] == 0) continue;
            if (d + 1 < per_depth.size() && per_depth[d + 1] > 0 && per_depth[d] + per_depth[d + 1] <= LANES) {
                r++;
                d++;
            } else {
                r += (per_depth[d] + LANES - 1) / LANES;
            }
        }
        return r;
    };
    // rounds on the longest path: the longest wave fold list, J's folds, the root fold
    // (counted as one round per two root children), J's backward list, the longest wave
    // backward list, and the two barriers
    auto evaluate = [&](const std::vector<int32_t> &owner) {
        int32_t wf = 0, wb = 0, jf = 0, jb = 0;
        for (int32_t o = -1; o < waves; o++) {
            std::vector<int32_t> cnt, bc(n, 0);
            fold_levels(owner, o, cnt);
            for (int32_t i = 1; i < n; i++)
                if (owner[i] == o) bc[depth[i]]++;
            if (o < 0) {
                jf = count_rounds(cnt);
                jb = bwd_rounds(bc);
            } else {
                wf = std::max(wf, count_rounds(cnt));
                wb = std::max(wb, bwd_rounds(bc));
            }
        }
        return wf + jf + (int32_t)(ch[0].size() + 1) / 2 + jb + wb;
    };
    // The Hines work goes to the first hw waves (hw = 1, 2, 4, .. waves, the cheapest); the
    // others have empty lists. With one Hines wave J is the root alone.
    std::vector<int32_t> best;
    int32_t best_cost = 0, best_hw = 1;
    for (int32_t hw = 1; hw <= waves; hw *= 2) {
        if (hw == 1) {
            std::vector<int32_t> owner(n, 0);
            owner[0] = -1;
            const int32_t cost = evaluate(owner);
            if (best.empty() || cost < best_cost) {
                best = owner;
                best_cost = cost;
                best_hw = 1;
            }
            continue;
        }
        const double share = (double)n / hw;
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
            std::vector<int64_t> load(hw, 0);
            load[0] += (int64_t)(first_bias * share);
            if (hw == waves) load[hw - 1] += (int64_t)(last_bias * share);
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
            const int32_t cost = evaluate(owner) + 2;   // two workgroup barriers
            if (cost < best_cost) {
                best = owner;
                best_cost = cost;
                best_hw = hw;
            }
        }
    }
    P.hines_waves = best_hw;
    P.longest = best_cost;
    const std::vector<int32_t> &owner = best;
    for (int32_t i = 0; i < n; i++) P.n_junction += owner[i] < 0 ? 1 : 0;

    // task lists: per wave, folds by level and backward tasks by depth; then J
    using Task = std::array<uint32_t, 4>;
    const size_t fold_nw = merge ? 4 : 2, align = fold_nw;
    std::vector<uint32_t> words;
    std::vector<std::pair<uint32_t, uint32_t>> rounds;   // (first word, count)
    std::vector<std::pair<uint32_t, uint32_t>> lists;    // (first round, round count)
    auto emit_list = [&](std::vector<std::vector<Task>> &levels, size_t nw) {
        const uint32_t r0 = (uint32_t)rounds.size();
        for (auto &tasks : levels)
            for (size_t k = 0; k < tasks.size(); k += LANES) {
                const size_t cnt = std::min(tasks.size() - k, (size_t)LANES);
                while (words.size() % align) words.push_back(0);   // rounds start at aligned indices
                rounds.emplace_back((uint32_t)words.size(), (uint32_t)cnt);
                for (size_t t = k; t < k + cnt; t++) words.insert(words.end(), tasks[t].begin(), tasks[t].begin() + (long)nw);
            }
        lists.emplace_back(r0, (uint32_t)rounds.size() - r0);
    };
    const uint32_t dummy = WAVE_NODE_BYTES * (uint32_t)n;
    auto off = [&](int32_t i) { return WAVE_NODE_BYTES * (uint32_t)P.pos[i]; };
    auto add_folds = [&](int32_t o, std::vector<std::vector<Task>> &f) {
        f.clear();
        fold_tasks(owner, o, [&](int32_t lv, int32_t i, int32_t k, bool m0, bool m1) {
            if ((int32_t)f.size() <= lv) f.resize(lv + 1);
            const size_t c = 2 * (size_t)k;
            const int32_t a = ch[i][c], b = c + 1 < ch[i].size() ? ch[i][c + 1] : -1;
            auto grand = [&](bool m, int32_t x, size_t g) {
                return m && g < ch[x].size() ? off(ch[x][g]) : dummy;
            };
            const uint32_t bo = b >= 0 ? off(b) : dummy;
            if (!merge) {
                f[lv].push_back({off(i) | (off(a) << 16), bo, 0u, 0u});
            } else {
                f[lv].push_back({off(i) | (off(a) << 16), bo | (grand(m0, a, 0) << 16),
                                 grand(m0, a, 1) | ((b >= 0 ? grand(m1, b, 0) : dummy) << 16),
                                 b >= 0 ? grand(m1, b, 1) : dummy});
            }
        });
        // position order within a level
        for (auto &t : f) std::stable_sort(t.begin(), t.end(), [](const Task &x, const Task &y) { return (x[0] & 0xFFFFu) < (y[0] & 0xFFFFu); });
    };
    // backward tasks by depth; two consecutive depths share a round when they have at most 64
    // tasks together: a node of the second depth whose parent is in the first solves that
    // parent again from the grandparent (the same operations on the same values) before
    // itself, so the round reads only values solved in earlier rounds
    auto add_bwd = [&](int32_t o, std::vector<std::vector<Task>> &b) {
        int32_t Dm = 0;
        for (int32_t i = 0; i < n; i++) Dm = std::max(Dm, depth[i]);
        std::vector<std::vector<int32_t>> lvl(Dm + 2);
        for (int32_t p = 0; p < n; p++) {
            const int32_t i = order[p];
            if (owner[i] == o && i > 0) lvl[depth[i]].push_back(i);
        }
        b.clear();
        auto one = [&](int32_t i) { return Task{off(i) | (off(parent[i]) << 16), dummy, 0u, 0u}; };
        for (int32_t d = 1; d <= Dm; d++) {
            if (lvl[d].empty()) continue;
            std::vector<Task> t;
            for (int32_t i : lvl[d]) t.push_back(one(i));
            if (!lvl[d + 1].empty() && lvl[d].size() + lvl[d + 1].size() <= (size_t)LANES) {
                for (int32_t i : lvl[d + 1]) {
                    const int32_t q = parent[i];
                    if (owner[q] == o)
                        t.push_back({off(i) | (off(parent[q]) << 16), off(q), 0u, 0u});
                    else
                        t.push_back(one(i));
                }
                d++;
            }
            b.push_back(t);
        }
    };
    for (int32_t w = 0; w <= waves; w++) {
        const int32_t o = w < waves ? w : -1;
        std::vector<std::vector<Task>> f, b;
        add_folds(o, f);
        emit_list(f, fold_nw);
        add_bwd(o, b);
        emit_list(b, 2);
    }
    while (words.size() % align) words.push_back(0);
    for (int k = 0; k < 5; k++) rounds.emplace_back((uint32_t)words.size(), 0u);   // sentinels
    const uint32_t nd = 2 * (uint32_t)waves + 2;
    uint32_t nh = (uint32_t)rounds.size();
    while ((nd + nh) % align) nh++;   // task words start at an aligned index
    for (auto &l : lists) P.block.push_back((nd + l.first) | (l.second << 16));
    for (auto &r : rounds) P.block.push_back((nd + nh + r.first) | (r.second << 16));
    while (P.block.size() < nd + nh) P.block.push_back(0u);
    P.block.insert(P.block.end(), words.begin(), words.end());
    P.block.insert(P.block.end(), (size_t)wave_plan_pad(merge), 0u);
    P.merged = merge;
    return P;
}

// Workgroup size for a cell of n compartments: each thread holds up to 4 compartments.
// limits (may be null): largest cell for 64, 128, 256 and 512 threads.
inline int32_t choose_tpb(int32_t n, int32_t forced, const int32_t *limits = nullptr) {
    if (forced == 64 || forced == 128 || forced == 256 || forced == 512)
        if (n <= forced * 4) return forced;
    static const int32_t dflt[4] = {256, 512, 1024, 2047};
    const bool custom = limits && (limits[0] | limits[1] | limits[2] | limits[3]) != 0;
    const int32_t *L = custom ? limits : dflt;
    for (int k = 0; k < 4; k++) {
        const int32_t tpb = 64 << k;
        if (n <= L[k] && n <= 4 * tpb) return tpb;
    }
    return 512;
}

}  // namespace edenhip
