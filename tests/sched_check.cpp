// Host check of the Hines schedules (eden_amd/csrc/tree_schedule.hpp): the wave-partitioned
// schedule of the tree kernel and the round schedule of the adjoint kernels.
//
// For random trees of several shapes and every workgroup size, the schedule block is decoded
// exactly as tree_kernel decodes it (round headers, task words, idle words aimed at the dummy
// slot n, the two-child straight-line fold plus the loop for further children) and the
// elimination is replayed round by round in fp32. The result must equal, bit for bit, the
// reference's serial elimination over the depth-first post-order with adjacency lists
// [parent, children ascending] (eden/Eden.cpp:4416-4486, 6114-6225). Structural checks:
// every node with children is folded exactly once, every non-root node is back-substituted
// exactly once, no round has more tasks than threads.
//
// Build: g++ -std=c++17 -O2 -ffp-contract=off -I<repo> sched_check.cpp; prints "ok <trials>".
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

#include "eden_amd/csrc/tree_schedule.hpp"

using namespace edenhip;

static std::vector<int32_t> make_parents(std::mt19937 &rng, int n, int shape) {
    std::vector<int32_t> par(n, -1);
    std::vector<int32_t> open{0}, nch(n, 0);
    for (int i = 1; i < n; i++) {
        switch (shape) {
        case 0: par[i] = (int32_t)(rng() % i); break;              // random recursive (any fan-out)
        case 1: {                                                    // binary, uniform open node
            const size_t k = rng() % open.size();
            par[i] = open[k];
            if (++nch[open[k]] == 2) {
                open[k] = open.back();
                open.pop_back();
            }
            open.push_back(i);
            break;
        }
        case 2: par[i] = i - 1; break;                               // unbranched cable
        case 3: par[i] = (int32_t)(rng() % std::min(i, 3)); break;  // star-like, wide fan-out
        case 4: par[i] = i > 8 ? (int32_t)(i - 1 - rng() % 8) : i - 1; break;   // long bushy chains
        case 5: par[i] = 0; break;                                   // star: every node on the root
        default: par[i] = i <= n / 2 ? i - 1 : n / 2; break;         // broom: chain, then a star
        }
    }
    return par;
}

// The wave-partitioned schedule (plan_tree_waves) decoded as tree_kernel decodes it:
// directory, per-wave lists of rounds (1..64 tasks), fold tasks of at most two children given
// by byte offsets (the dummy slot n as the missing second child), the junction lists run by
// wave 0 between the two workgroup barriers with the root fold (all root children in order)
// and the root solve in between. Within a round all tasks read before any writes (one SIMD
// instruction stream); rounds of a wave follow in order; waves are replayed one after the
// other in reverse order, and a read of a node that another wave writes in the same phase is
// an error (so no order of the waves could change the result). Every node must receive its
// children in ascending order, each exactly once.
//
// With merge (the constant-tree form), the fold lists are replayed twice, as the kernel runs
// them: once for the eliminated diagonals (the set-up pre-sweep), then, with the ratios
// Uj / D' fixed, for the right-hand sides; a merged child's value is formed again from the
// grandchildren in the parent's task, from the values at the round's start.
static int check_waves(int n, const std::vector<int32_t> &par, const std::vector<std::vector<int>> &ch,
                       const std::vector<float> &D0, const std::vector<float> &X0, const std::vector<float> &ui,
                       const std::vector<float> &uj, const std::vector<float> &Dr, const std::vector<float> &Xr,
                       int waves, int trial, int shape, bool merge) {
    const WavePlan P = plan_tree_waves(n, par.data(), waves, 0.0, 0.0, merge);
    const std::vector<uint32_t> &B = P.block;
    int bad = P.merged == merge ? 0 : 1;
    const int fnw = merge ? 4 : 2;        // words per fold task; rounds start at multiples of it
    const int reach = 64 * fnw;           // words a wave may read from any round's start
    std::vector<float> dx(n + 1), dv(n + 1), su(n + 1), sj(n + 1), rat(n + 1, 1.0f);
    for (int i = 0; i < n; i++) {
        dx[P.pos[i]] = D0[i];
        dv[P.pos[i]] = X0[i];
        su[P.pos[i]] = i ? ui[i] : 0.0f;
        sj[P.pos[i]] = i ? uj[i] : 0.0f;
    }
    dx[n] = 1.0f; dv[n] = 0.0f; su[n] = 0.0f; sj[n] = 0.0f;
    std::vector<std::vector<int>> got(n + 1);   // children folded into each position, in order
    std::vector<int> solved(n + 1, 0), writer(n + 1, -1);
    const int nd = 2 * waves + 2;
    auto slot = [&](uint32_t off, bool &ok) {
        if (off % WAVE_NODE_BYTES != 0 || off / WAVE_NODE_BYTES > (uint32_t)n) ok = false;
        return ok ? (int)(off / WAVE_NODE_BYTES) : n;
    };
    auto fold = [&](int p, int c) {   // one child into p, the reference's operations
        const float r = sj[c] / dx[c];
        dx[p] = dx[p] - r * su[c];
        dv[p] = dv[p] - r * dv[c];
        if (c != n) got[p].push_back(c);
    };
    enum Kind { FOLD, DIAG, RHS, BWD };
    auto run = [&](int dir, int wave, Kind kind, bool owner_check) -> bool {
        if (dir >= (int)B.size()) return false;
        const int first = B[dir] & 0xFFFF, cnt = B[dir] >> 16;
        if (cnt > 0 && (first < nd || first + cnt + 5 > (int)B.size())) return false;
        const int nw = kind == BWD || !merge ? 2 : 4;
        for (int r = 0; r < cnt; r++) {
            const uint32_t h = B[first + r];
            const int off = h & 0xFFFF, tc = h >> 16;
            // every lane reads a task from the round: within the block, aligned
            if (tc < 1 || tc > 64 || off + reach > (int)B.size() || (off % fnw)) return false;
            // reads of the whole round first
            struct T { int p, a, b; float ox, oy; };
            std::vector<T> tasks;
            std::vector<int> two_step;
            bool ok = true;
            for (int t = 0; t < tc; t++) {
                auto touch = [&](int x) {
                    if (owner_check && x != n && writer[x] >= 0 && writer[x] != wave) bad++;
                };
                const uint32_t *w = &B[off + nw * t];
                if (kind == FOLD) {
                    const int p = slot(w[0] & 0xFFFF, ok), a = slot(w[0] >> 16, ok), b = slot(w[1], ok);
                    if (!ok || p == n || a == n || p == 0) return false;
                    touch(p); touch(a); touch(b);
                    // the kernel's two-child form: ((D - r0 Ui0) - r1 Ui1), dummy b gives 0
                    const float r0 = sj[a] / dx[a], r1 = sj[b] / dx[b];
                    tasks.push_back({p, a, b, (dx[p] - r0 * su[a]) - r1 * su[b], (dv[p] - r0 * dv[a]) - r1 * dv[b]});
                } else if (kind == DIAG || kind == RHS) {
                    const int p = slot(w[0] & 0xFFFF, ok), a = slot(w[0] >> 16, ok), b = slot(w[1] & 0xFFFF, ok);
                    const int a0 = slot(w[1] >> 16, ok), a1 = slot(w[2] & 0xFFFF, ok);
                    const int b0 = slot(w[2] >> 16, ok), b1 = slot(w[3], ok);
                    if (!ok || p == n || a == n || p == 0) return false;
                    for (int x : {p, a, b, a0, a1, b0, b1}) touch(x);
                    if (kind == DIAG) {
                        auto dprime = [&](int c, int g0, int g1) {
                            return (dx[c] - (sj[g0] / dx[g0]) * su[g0]) - (sj[g1] / dx[g1]) * su[g1];
                        };
                        const float da = dprime(a, a0, a1), db = dprime(b, b0, b1);
                        tasks.push_back({p, a, b, (dx[p] - (sj[a] / da) * su[a]) - (sj[b] / db) * su[b], dv[p]});
                    } else {
                        auto rhs = [&](int c, int g0, int g1) { return (dv[c] - rat[g0] * dv[g0]) - rat[g1] * dv[g1]; };
                        const float xa = rhs(a, a0, a1), xb = rhs(b, b0, b1);
                        tasks.push_back({p, a, b, dx[p], (dv[p] - rat[a] * xa) - rat[b] * xb});
                    }
                } else {
                    const int p = slot(w[0] & 0xFFFF, ok), sp = slot(w[0] >> 16, ok), m = slot(w[1], ok);
                    if (!ok || p == n || sp == n || p == 0) return false;
                    touch(p); touch(sp); touch(m);
                    if (solved[sp] == 0) bad++;   // the solved node it starts from is not solved yet
                    float xm = dv[sp];
                    if (m != n) {
                        // the parent solved again: it must be unsolved at the round's start, and
                        // its own task in this round (checked after the round)
                        if (solved[m] != 0) bad++;
                        two_step.push_back(m);
                        xm = (dv[m] - su[m] * dv[sp]) / dx[m];
                    }
                    tasks.push_back({p, sp, -1, dx[p], (dv[p] - su[p] * xm) / dx[p]});
                }
            }
            for (auto &t : tasks) {
                dx[t.p] = t.ox;
                dv[t.p] = t.oy;
                if (kind == FOLD || kind == RHS) {
                    got[t.p].push_back(t.a);
                    if (t.b != n) got[t.p].push_back(t.b);
                } else if (kind == BWD) {
                    solved[t.p]++;
                }
            }
            for (int m : two_step)
                if (solved[m] != 1) bad++;   // the parent's own task was not in the round
        }
        return true;
    };
    bool ok = (int)B.size() >= nd + 1;
    // every header a wave may read (its lists' and five past them) points at an aligned
    // offset with a full round's reach of block after it
    for (int d = 0; d < nd && ok; d++) {
        const int first = B[d] & 0xFFFF, cnt = B[d] >> 16;
        for (int r = 0; r < cnt + 5 && ok; r++) {
            if (first + r >= (int)B.size()) { ok = false; break; }
            const int off = B[first + r] & 0xFFFF;
            if ((off % fnw) || off + reach > (int)B.size()) ok = false;
        }
    }
    // the nodes each wave writes in a phase, from its list (before any replay)
    auto owners = [&](int odd) {
        const int nw = odd ? 2 : fnw;
        std::fill(writer.begin(), writer.end(), -1);
        for (int w = 0; w < waves && ok; w++) {
            const int first = B[2 * w + odd] & 0xFFFF, cnt = B[2 * w + odd] >> 16;
            for (int r = 0; r < cnt && ok; r++) {
                const uint32_t h = B[first + r];
                for (int t = 0; t < (int)(h >> 16); t++) {
                    const int p = (int)((B[(h & 0xFFFF) + nw * t] & 0xFFFF) / WAVE_NODE_BYTES);
                    if (p > n || (writer[p] >= 0 && writer[p] != w)) bad++;
                    writer[p] = w;
                }
            }
        }
    };
    owners(0);
    if (merge) {
        // set-up pre-sweep: eliminated diagonals below the root, then the fixed ratios
        for (int w = waves - 1; w >= 0 && ok; w--) ok = run(2 * w, w, DIAG, true);
        std::vector<int> keep = writer;
        std::fill(writer.begin(), writer.end(), -1);
        if (ok) ok = run(2 * waves, 0, DIAG, false);
        writer = keep;
        for (int i = 1; i < n; i++) rat[P.pos[i]] = sj[P.pos[i]] / dx[P.pos[i]];
    }
    for (int w = waves - 1; w >= 0 && ok; w--) ok = run(2 * w, w, merge ? RHS : FOLD, true);
    std::fill(writer.begin(), writer.end(), -1);
    if (ok) ok = run(2 * waves, 0, merge ? RHS : FOLD, false);
    // the root fold: every child of the root (positions 1 .. nc) in order, then the solve
    for (size_t k = 0; k < ch[0].size() && ok; k++) {
        const int c = P.pos[ch[0][k]];
        if (merge) {
            dx[0] = dx[0] - rat[c] * su[c];
            dv[0] = dv[0] - rat[c] * dv[c];
            got[0].push_back(c);
        } else {
            fold(0, c);
        }
    }
    dv[0] = dv[0] / dx[0];
    solved[0] = 1;
    if (ok) ok = run(2 * waves + 1, 0, BWD, false);
    if (ok) owners(1);
    for (int w = waves - 1; w >= 0 && ok; w--) ok = run(2 * w + 1, w, BWD, true);
    if (!ok) {
        std::printf("trial %d shape %d n %d waves %d merge %d: malformed block\n", trial, shape, n, waves, merge);
        return bad + 1;
    }
    for (int i = 0; i < n; i++) {
        const int p = P.pos[i];
        std::vector<int> want;
        for (int c : ch[i]) want.push_back(P.pos[c]);
        if (got[p] != want) {
            std::printf("trial %d shape %d n %d waves %d merge %d: node %d children folded out of order\n", trial,
                        shape, n, waves, merge, i);
            bad++;
        }
        if (solved[p] != 1) bad++;
        if (std::memcmp(&dv[p], &Xr[i], 4) != 0 || std::memcmp(&dx[p], &Dr[i], 4) != 0) {
            std::printf("trial %d shape %d n %d waves %d merge %d: node %d gives %.9g, serial %.9g\n", trial, shape,
                        n, waves, merge, i, dv[p], Xr[i]);
            bad++;
            break;
        }
    }
    if (dx[n] != 1.0f || dv[n] != 0.0f) bad++;
    // waves beyond the Hines waves have empty lists
    if (P.hines_waves < 1 || P.hines_waves > waves) bad++;
    for (int w = P.hines_waves; w < waves; w++)
        if ((B[2 * w] >> 16) != 0 || (B[2 * w + 1] >> 16) != 0) bad++;
    return bad;
}

int main(int argc, char **argv) {
    const int trials = argc > 1 ? std::atoi(argv[1]) : 200;
    std::mt19937 rng(12345);
    std::uniform_real_distribution<float> U(0.5f, 2.0f);
    int bad = 0, runs = 0;
    for (int trial = 0; trial < trials && bad == 0; trial++) {
        const int n = trial % 9 == 0 ? 2047 : 2 + (int)(rng() % (trial % 4 == 0 ? 2046 : 300));
        const int shape = trial % 7;
        const std::vector<int32_t> par = make_parents(rng, n, shape);
        std::vector<float> D0(n), X0(n), ui(n), uj(n);
        for (int i = 0; i < n; i++) {
            D0[i] = 1.0f + 3.0f * U(rng);
            X0[i] = U(rng) - 1.25f;
            ui[i] = -U(rng) * 0.3f;
            uj[i] = -U(rng) * 0.3f;
        }
        // reference: serial elimination over the DFS post-order, then back-substitution
        std::vector<std::vector<int>> ch(n);
        for (int i = 1; i < n; i++) ch[par[i]].push_back(i);
        std::vector<int> order;
        std::vector<std::pair<int, size_t>> st{{0, 0}};
        while (!st.empty()) {
            auto &[v, k] = st.back();
            if (k < ch[v].size()) {
                const int c = ch[v][k++];
                st.push_back({c, 0});
            } else {
                order.push_back(v);
                st.pop_back();
            }
        }
        std::vector<float> Dr = D0, Xr = X0;
        for (int k = 0; k < n - 1; k++) {
            const int i = order[k], j = par[i];
            const float r = uj[i] / Dr[i];
            Dr[j] -= r * ui[i];
            Xr[j] -= r * Xr[i];
        }
        Xr[0] = Xr[0] / Dr[0];
        for (int k = n - 2; k >= 0; k--) {
            const int i = order[k], j = par[i];
            Xr[i] = (Xr[i] - ui[i] * Xr[j]) / Dr[i];
        }
        for (int waves : {1, 2, 4, 8})
            for (bool merge : {false, true}) {
                runs++;
                bad += check_waves(n, par, ch, D0, X0, ui, uj, Dr, Xr, waves, trial, shape, merge);
            }
        for (int tpb : {64, 128, 256, 512}) {
            runs++;
            const TreePlan P = plan_tree(n, par.data(), tpb);
            // the round schedule's 10-bit child count: such trees are flagged, not planned
            size_t widest = 0;
            for (auto &c : ch) widest = std::max(widest, c.size());
            if (!P.ok) {
                if (widest <= 1023) bad++;
                continue;
            }
            if (widest > 1023) bad++;
            const uint32_t *B = P.block.data();
            const int RF = P.rounds_fwd, RB = P.rounds_bwd;
            // LDS image by position, with the dummy slot n
            std::vector<float> dx(n + 1), dv(n + 1), su(n + 1), sj(n + 1);
            for (int i = 0; i < n; i++) {
                dx[P.pos[i]] = D0[i];
                dv[P.pos[i]] = X0[i];
                su[P.pos[i]] = i ? ui[i] : 0.0f;
                sj[P.pos[i]] = i ? uj[i] : 0.0f;
            }
            dx[n] = 1.0f; dv[n] = 0.0f; su[n] = 0.0f; sj[n] = 0.0f;
            std::vector<int> folded(n + 1, 0), solved(n + 1, 0);
            auto word = [&](int r, int t, uint32_t idle) {
                const uint32_t h = B[r];
                return (uint32_t)t < (h >> 16) ? B[(h & 0xFFFFu) + t] : idle;
            };
            for (int r = 0; r < RF + RB + 2; r++)
                if ((int)(B[r] >> 16) > tpb) bad++;
            for (int r = 0; r < RF; r++) {
                // all tasks of a round read before any writes (distinct heights do not mix)
                std::vector<std::pair<int, std::pair<float, float>>> out;
                for (int t = 0; t < tpb; t++) {
                    const uint32_t w = word(r, t, idle_fwd_word(n));
                    const int p = w & 0x7FF, fc = (w >> 11) & 0x7FF, nc = (int)(w >> 22);
                    const int c1 = nc >= 2 ? fc + 1 : n;
                    const float r0 = sj[fc] / dx[fc];
                    const float r1 = sj[c1] / dx[c1];
                    float ox = (dx[p] - r0 * su[fc]) - r1 * su[c1];
                    float oy = (dv[p] - r0 * dv[fc]) - r1 * dv[c1];
                    for (int i = 2; i < nc; i++) {
                        const float ratio = sj[fc + i] / dx[fc + i];
                        ox -= ratio * su[fc + i];
                        oy -= ratio * dv[fc + i];
                    }
                    if (p != n) folded[p]++;
                    out.push_back({p, {ox, oy}});
                }
                for (auto &o : out) {
                    dx[o.first] = o.second.first;
                    dv[o.first] = o.second.second;
                }
            }
            dv[0] = dv[0] / dx[0];
            for (int r = 0; r < RB; r++) {
                std::vector<std::pair<int, float>> out;
                for (int t = 0; t < tpb; t++) {
                    const uint32_t w = word(RF + 1 + r, t, idle_bwd_word(n));
                    const int p = w & 0x7FF, pp = (w >> 11) & 0x7FF;
                    if (p != n) solved[p]++;
                    out.push_back({p, (dv[p] - su[p] * dv[pp]) / dx[p]});
                }
                for (auto &o : out) dv[o.first] = o.second;
            }
            if (dx[n] != 1.0f || dv[n] != 0.0f) bad++;
            for (int i = 0; i < n; i++) {
                const int p = P.pos[i];
                if (folded[p] != (ch[i].empty() ? 0 : 1)) bad++;
                if (solved[p] != (i ? 1 : 0)) bad++;
                if (std::memcmp(&dv[p], &Xr[i], 4) != 0 || std::memcmp(&dx[p], &Dr[i], 4) != 0) {
                    std::printf("trial %d shape %d n %d tpb %d: node %d gives %.9g, serial %.9g\n", trial, shape, n,
                                tpb, i, dv[p], Xr[i]);
                    bad++;
                    break;
                }
            }
        }
    }
    if (bad) {
        std::printf("FAIL %d\n", bad);
        return 1;
    }
    std::printf("ok %d\n", runs);
    return 0;
}
