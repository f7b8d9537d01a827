// Host-side tables for the coupling operator: mode indexing, the Wigner-3j
// translation table ab5, Legendre recurrence coefficients and the pair-GEMM
// work plan with its ab5 contraction lists.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "internal.hpp"

namespace yasfhip {

// Mode order n = (tau-1) L(L+2) + l^2 - 1 + l + m
// (reference yasfpy/functions/cpu_numba.py:395-405, misc.py:23-43).
std::vector<Mode> modes_of(int lmax) {
    std::vector<Mode> out;
    for (int tau = 1; tau <= 2; ++tau)
        for (int l = 1; l <= lmax; ++l)
            for (int m = -l; m <= l; ++m)
                out.push_back({tau, l, m});
    return out;
}

namespace {

long double factorial(int n) {
    static std::vector<long double> table = [] {
        std::vector<long double> t(80);
        t[0] = 1.0L;
        for (int i = 1; i < 80; ++i)
            t[i] = t[i - 1] * static_cast<long double>(i);
        return t;
    }();
    if (n < 0 || n >= 80)
        throw Error(kInternal, "factorial argument out of range");
    return table[n];
}

// Wigner 3j symbol by Racah's single-sum formula, evaluated in long double.
// The alternating sum is accumulated from its largest-index term downwards.
long double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (m1 + m2 + m3 != 0)
        return 0.0L;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3)
        return 0.0L;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2)
        return 0.0L;
    const int kmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    const int kmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    long double sum = 0.0L, mag = 0.0L;
    for (int k = kmax; k >= kmin; --k) {
        const long double den = factorial(k) * factorial(j3 - j2 + k + m1) *
                                factorial(j3 - j1 + k - m2) * factorial(j1 + j2 - j3 - k) *
                                factorial(j1 - k - m1) * factorial(j2 - k + m2);
        sum += ((k & 1) ? -1.0L : 1.0L) / den;
        mag += 1.0L / den;
    }
    // Non-trivial zeros of the 3j symbol cancel exactly in exact arithmetic;
    // in long double they leave a residue of a few ulps of the term sum.
    if (std::fabs(sum) <= 256.0L * LDBL_EPSILON * mag)
        return 0.0L;
    const long double delta = factorial(j1 + j2 - j3) * factorial(j1 - j2 + j3) *
                              factorial(-j1 + j2 + j3) / factorial(j1 + j2 + j3 + 1);
    const long double pref = factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) *
                             factorial(j2 - m2) * factorial(j3 + m3) * factorial(j3 - m3);
    const int e = j1 - j2 - m3;
    const long double sign = (((e % 2) + 2) % 2) ? -1.0L : 1.0L;
    return sign * std::sqrt(delta * pref) * sum;
}

}  // namespace

// ab5[j1, j2, p] for (tau1,l1,m1) = modes[j1], (tau2,l2,m2) = modes[j2]
// (reference yasfpy/numerics.py:224-272).  i^k is applied exactly.
std::vector<cplx> translation_table(int lmax) {
    const auto modes = modes_of(lmax);
    const int nmax = static_cast<int>(modes.size());
    const int P = 2 * lmax + 1;
    std::vector<cplx> ab5(static_cast<size_t>(nmax) * nmax * P, cplx(0.0, 0.0));
    const cplx ipow[4] = {cplx(1, 0), cplx(0, 1), cplx(-1, 0), cplx(0, -1)};
    for (int j1 = 0; j1 < nmax; ++j1) {
        const auto [tau1, l1, m1] = modes[j1];
        for (int j2 = 0; j2 < nmax; ++j2) {
            const auto [tau2, l2, m2] = modes[j2];
            for (int p = 0; p < P; ++p) {
                const int e = std::abs(m1 - m2) - std::abs(m1) - std::abs(m2) + l2 - l1 + p;
                const cplx phase = ipow[((e % 4) + 4) % 4];
                const double sgn = ((m1 - m2) % 2 != 0) ? -1.0 : 1.0;
                const double norm = std::sqrt(static_cast<double>((2 * l1 + 1) * (2 * l2 + 1)) /
                                              (2.0 * l1 * (l1 + 1) * l2 * (l2 + 1)));
                double mag = 0.0;
                if (tau1 == tau2) {
                    const long double wa = wigner3j(l1, l2, p, m1, -m2, -m1 + m2);
                    const long double wb = wigner3j(l1, l2, p, 0, 0, 0);
                    mag = static_cast<double>(l1 * (l1 + 1) + l2 * (l2 + 1) - p * (p + 1)) *
                          std::sqrt(2.0 * p + 1.0) * static_cast<double>(wa * wb);
                } else if (p > 0) {
                    const long rad = static_cast<long>(l1 + l2 + 1 + p) * (l1 + l2 + 1 - p) *
                                     (p + l1 - l2) * (p - l1 + l2) * (2 * p + 1);
                    const long double wa = wigner3j(l1, l2, p, m1, -m2, -m1 + m2);
                    const long double wb = wigner3j(l1, l2, p - 1, 0, 0, 0);
                    // A negative radicand only occurs where a 3j factor vanishes.
                    if (rad > 0)
                        mag = std::sqrt(static_cast<double>(rad)) * static_cast<double>(wa * wb);
                }
                ab5[(static_cast<size_t>(j1) * nmax + j2) * P + p] = phase * (sgn * norm * mag);
            }
        }
    }
    return ab5;
}

// Work plan of the factorised coupling matvec
//   Q[s1][p][dm][n2] = sum_{s2 != s1} Ptilde_p^{|dm|}(cos th) h_p(k d) e^{i dm phi} x[s2][n2]
//   (W x)[s1][n1]   = sum_{p, n2} ab5[n2][n1][p] Q[s1][p][m2-m1][n2]
// which regroups the reference's quadruple sum (cpu_numba.py:50-93) so the
// pair-independent ab5 contraction leaves the O(N^2) loop.
namespace {

// Process-wide caches: ab5 and the plan depend on lmax only, and take up to
// 0.12 s and 0.23 s to build at lmax 8 -- once per process instead of once
// per session, preconditioner build or server.
std::mutex g_cache_mutex;
std::unique_ptr<const std::vector<cplx>> g_ab5[kMaxLmax + 1];
std::unique_ptr<const CouplingPlan> g_plans[kMaxLmax + 1];

}  // namespace

const std::vector<cplx> &translation_table_cached(int lmax) {
    YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    if (!g_ab5[lmax])
        g_ab5[lmax] = std::make_unique<const std::vector<cplx>>(translation_table(lmax));
    return *g_ab5[lmax];
}

const CouplingPlan &coupling_plan_cached(int lmax) {
    YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
    {
        std::lock_guard<std::mutex> lock(g_cache_mutex);
        if (g_plans[lmax])
            return *g_plans[lmax];
    }
    auto plan = std::make_unique<const CouplingPlan>(build_coupling_plan(lmax));  // outside the lock
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    if (!g_plans[lmax])
        g_plans[lmax] = std::move(plan);
    return *g_plans[lmax];
}

CouplingPlan build_coupling_plan(int lmax) {
    YH_REQUIRE(lmax >= 1 && lmax <= kMaxLmax, "lmax must be in [1, 8]");
    CouplingPlan plan;
    plan.lmax = lmax;
    plan.nmax = nmax_of(lmax);
    plan.ndm = 4 * lmax + 1;
    const auto modes = modes_of(lmax);
    const int P = 2 * lmax + 1;

    // Source columns compatible with each signed dm: |m2 - dm| <= L, ordered
    // by decreasing l2.  Order p needs only columns with l2 >= p - L
    // (triangle rule of the 3j symbols), so with this order the columns an
    // order needs form a prefix that shrinks as p grows.
    std::vector<std::vector<int>> cols(plan.ndm);
    for (int dmi = 0; dmi < plan.ndm; ++dmi) {
        const int dm = dmi - 2 * lmax;
        for (int n2 = 0; n2 < plan.nmax; ++n2)
            if (std::abs(modes[n2].m - dm) <= lmax)
                cols[dmi].push_back(n2);
        std::stable_sort(cols[dmi].begin(), cols[dmi].end(),
                         [&](int a, int b) { return modes[a].l > modes[b].l; });
    }
    // need[dmi][p]: column tiles of the dm that order p contributes to (one
    // past the last tile holding a column with a nonzero ab5 term at p).
    const auto &ab5 = translation_table_cached(lmax);
    std::vector<std::vector<int>> need(plan.ndm, std::vector<int>(P, 0));
    for (int dmi = 0; dmi < plan.ndm; ++dmi) {
        const int dm = dmi - 2 * lmax;
        for (size_t c = 0; c < cols[dmi].size(); ++c) {
            const int n2 = cols[dmi][c];
            for (int n1 = 0; n1 < plan.nmax; ++n1) {
                if (modes[n2].m - modes[n1].m != dm)
                    continue;
                for (int p = 0; p < P; ++p)
                    if (ab5[(static_cast<size_t>(n2) * plan.nmax + n1) * P + p] != cplx(0.0, 0.0))
                        need[dmi][p] = std::max(need[dmi][p], static_cast<int>(c) / 16 + 1);
            }
        }
    }
    plan.order_tiles.assign(static_cast<size_t>(plan.ndm) * P, 0);
    for (int dmi = 0; dmi < plan.ndm; ++dmi)
        for (int p = 0; p < P; ++p)
            plan.order_tiles[static_cast<size_t>(dmi) * P + p] = need[dmi][p];
    // orders a job over column tiles [t0, ...) must compute: the prefix of
    // p = |dm|, |dm|+1, ... up to the last order that needs a tile >= t0
    auto orders_for = [&](int dmi, int t0) {
        const int adm = std::abs(dmi - 2 * lmax);
        int nord = 0;
        for (int p = adm; p < P; ++p)
            if (need[dmi][p] > t0)
                nord = p - adm + 1;
        return nord;
    };
    int maxcols = 0;
    for (auto &c : cols)
        maxcols = std::max<int>(maxcols, ((static_cast<int>(c.size()) + 15) / 16) * 16);
    plan.maxcols = maxcols;
    plan.colmap.assign(static_cast<size_t>(plan.ndm) * maxcols, -1);
    for (int dmi = 0; dmi < plan.ndm; ++dmi)
        for (size_t c = 0; c < cols[dmi].size(); ++c)
            plan.colmap[static_cast<size_t>(dmi) * maxcols + c] = cols[dmi][c];

    // Q slots: per dm, R orders x all of its column tiles, slot(ip, it) =
    // qbase(dm) + ip * ntiles(dm) + it; slot[(dmi, p, col)] -> Q column index
    // slot * 16 + col % 16.  Jobs: (dm, all its orders, a group of column
    // tiles) with at most kJobMaxTiles C tiles for the on-the-fly kernel; the
    // table GEMM's lists group the same column tiles more widely.
    std::map<std::tuple<int, int, int>, int> slot;
    int q = 0;
    long long macs = 0;
    for (int dmi = 0; dmi < plan.ndm; ++dmi) {
        const int dm = dmi - 2 * lmax;
        const int R = P - std::abs(dm);
        const int ntiles = (static_cast<int>(cols[dmi].size()) + 15) / 16;
        for (int ip = 0; ip < R; ++ip)
            for (int it = 0; it < ntiles; ++it)
                for (int c = 0; c < 16; ++c)
                    slot[{dmi, std::abs(dm) + ip, it * 16 + c}] = (q + ip * ntiles + it) * 16 + c;
        const int group = std::max(1, std::min(kJobMaxColTiles, kJobMaxTiles / R));
        for (int t0 = 0; t0 < ntiles; t0 += group) {
            const int nt = std::min(group, ntiles - t0);
            const int nord = orders_for(dmi, t0);
            if (nord == 0)
                continue;
            plan.jobs.push_back(PairJob{dm, t0, nt, q, nord * nt, ntiles, nord, 0});
            plan.stage_elems = std::max(plan.stage_elems, R * 256 + 16 * nt * 16);
            macs += static_cast<long long>(nord) * nt * 16;
        }
        for (int g = 0; g < kTableGroupings; ++g) {
            const int njob = (ntiles + kTableGrouping[g] - 1) / kTableGrouping[g];
            for (int j = 0, t0 = 0; j < njob; ++j) {  // even split: sizes differ by at most one
                const int nt = ntiles / njob + (j < ntiles % njob ? 1 : 0);
                const int nord = orders_for(dmi, t0);
                if (nord > 0)
                    plan.table_jobs[g].push_back(PairJob{dm, t0, nt, q, nord * nt, ntiles, nord, 0});
                t0 += nt;
            }
        }
        {
            const int njob = (ntiles + kTableWideTiles64 - 1) / kTableWideTiles64;
            for (int j = 0, t0 = 0; j < njob; ++j) {
                const int nt = ntiles / njob + (j < ntiles % njob ? 1 : 0);
                const int nord = orders_for(dmi, t0);
                if (nord > 0)
                    plan.table_jobs_w64.push_back(PairJob{dm, t0, nt, q, nord * nt, ntiles, nord, 0});
                t0 += nt;
            }
        }
        q += R * ntiles;
    }
    plan.nq = q;
    plan.pair_macs_per_pair = macs;
    auto by_cost = [](const PairJob &a, const PairJob &b) { return a.cost > b.cost; };
    std::stable_sort(plan.jobs.begin(), plan.jobs.end(), by_cost);
    for (auto &tj : plan.table_jobs)
        std::stable_sort(tj.begin(), tj.end(), by_cost);
    std::stable_sort(plan.table_jobs_w64.begin(), plan.table_jobs_w64.end(), by_cost);
    // fp16 GEMM bins.  Jobs of the grouping that matches the capacity are
    // cut into sub-jobs of at most 8 (two-term) or 16 (one-term, two per
    // wave) orders, and the sub-jobs are packed first-fit, widest first, into
    // bins of kHalfBinWaves waves and cap column tiles.  A sub-job of w waves
    // deals its orders ip0 + i and ip0 + i + w to its wave i.
    for (int form = 0; form < 2; ++form) {
        const int nip = form + 1;
        for (int c = 0; c < kHalfBinCaps; ++c) {
            const int cap = kHalfBinCap[c];
            int g = 0;
            while (kTableGrouping[g] != cap)
                ++g;
            struct Sub {
                const PairJob *j;
                int ip0, n, w;
            };
            std::vector<Sub> subs;
            for (const PairJob &j : plan.table_jobs[g])
                for (int ip0 = 0; ip0 < j.nord; ip0 += kHalfBinWaves * nip) {
                    const int n = std::min(kHalfBinWaves * nip, j.nord - ip0);
                    subs.push_back(Sub{&j, ip0, n, (n + nip - 1) / nip});
                }
            std::stable_sort(subs.begin(), subs.end(), [](const Sub &a, const Sub &b) {
                return a.j->ntile != b.j->ntile ? a.j->ntile > b.j->ntile : a.w > b.w;
            });
            std::vector<HalfBin> bins;
            std::vector<int> waves_used, widest;
            for (const Sub &sj : subs) {
                const int nt = sj.j->ntile;
                size_t bi = 0;
                while (bi < bins.size() &&
                       (waves_used[bi] + sj.w > kHalfBinWaves || bins[bi].ntile + nt > cap))
                    ++bi;
                if (bi == bins.size()) {
                    HalfBin nb{};
                    std::fill(std::begin(nb.tcol), std::end(nb.tcol), -1);
                    std::fill(std::begin(nb.wd), std::end(nb.wd), -1);
                    for (int w = 0; w < kHalfBinWaves; ++w)
                        nb.wip[w][0] = nb.wip[w][1] = nb.wq[w][0] = nb.wq[w][1] = -1;
                    bins.push_back(nb);
                    waves_used.push_back(0);
                    widest.push_back(0);
                }
                HalfBin &b = bins[bi];
                const int dmi = sj.j->dm + 2 * lmax, tb = b.ntile;
                for (int k = 0; k < nt; ++k)
                    b.tcol[tb + k] = dmi * maxcols + (sj.j->tile0 + k) * 16;
                for (int i = 0; i < sj.w; ++i) {
                    const int w = waves_used[bi] + i;
                    b.wd[w] = dmi;
                    b.wtb[w] = tb;
                    b.wte[w] = tb + nt;
                    for (int o = 0; o < nip; ++o) {
                        const int io = i + o * sj.w;
                        if (io >= sj.n)
                            continue;
                        const int ip = sj.ip0 + io;
                        b.wip[w][o] = ip;
                        b.wq[w][o] = sj.j->qbase + ip * sj.j->tstride + sj.j->tile0 - tb;
                    }
                }
                waves_used[bi] += sj.w;
                b.ntile += nt;
                widest[bi] = std::max(widest[bi], nt);
            }
            // widest bins first (the longest blocks start first)
            std::vector<size_t> order(bins.size());
            for (size_t i = 0; i < order.size(); ++i)
                order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return widest[a] > widest[b]; });
            int tiles = 0;
            for (size_t i : order) {
                plan.half_bins[form][c].push_back(bins[i]);
                tiles += widest[i];
            }
            plan.half_bin_tiles[form][c] = tiles;
            // every (order, column tile) product of the grouping's jobs
            // exactly once, with the column tile the job gives it
            std::vector<int> want(static_cast<size_t>(q), -1), seen(static_cast<size_t>(q), 0);
            for (const PairJob &j : plan.table_jobs[g])
                for (int ip = 0; ip < j.nord; ++ip)
                    for (int k = 0; k < j.ntile; ++k)
                        want[j.qbase + ip * j.tstride + j.tile0 + k] = (j.dm + 2 * lmax) * maxcols + (j.tile0 + k) * 16;
            for (const HalfBin &b : plan.half_bins[form][c])
                for (int w = 0; w < kHalfBinWaves; ++w)
                    for (int o = 0; o < 2; ++o) {
                        if (b.wd[w] < 0 || b.wip[w][o] < 0)
                            continue;
                        for (int k = b.wtb[w]; k < b.wte[w]; ++k) {
                            const int slot = b.wq[w][o] + k;
                            if (slot < 0 || slot >= q || want[slot] != b.tcol[k] ||
                                b.tcol[k] / maxcols != b.wd[w])
                                throw Error(kInternal, "fp16 GEMM bin outside the pair plan");
                            ++seen[slot];
                        }
                    }
            for (int slot = 0; slot < q; ++slot)
                if (seen[slot] != (want[slot] >= 0 ? 1 : 0))
                    throw Error(kInternal, "fp16 GEMM bins do not cover the pair plan once");
        }
    }

    // Contraction lists: for each target mode n1, the nonzero ab5[n2][n1][p]
    // in (n2, p) order, as rows of a CSR.  A Q slot (dm, p, column n2) feeds
    // only the modes with m1 = m2 - dm, so the rows are ordered to keep each
    // m1's modes inside one block of kContractModes rows: the m1 sets, largest
    // first, are packed first-fit into contract_groups(nmax) blocks (padding
    // rows: mode -1), and each Q row is then read by one block.
    // crow[k], crow[k + 1]: row k's entries; crow[R + 1 + k]: its mode
    // (R = contract_groups(nmax) kContractModes rows).
    const int crows = contract_groups(plan.nmax) * kContractModes;
    std::vector<int> row_mode(crows, -1);
    {
        std::vector<std::pair<int, int>> sets;  // (mode count, m1)
        for (int m = -lmax; m <= lmax; ++m) {
            int c = 0;
            for (const auto &md : modes)
                c += md.m == m;
            sets.push_back({c, m});
        }
        std::stable_sort(sets.begin(), sets.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        std::vector<std::vector<int>> groups;
        for (const auto &st : sets) {
            size_t g = 0;
            while (g < groups.size() && groups[g].size() + st.first > static_cast<size_t>(kContractModes))
                ++g;
            if (g == groups.size())
                groups.emplace_back();
            for (int n1 = 0; n1 < plan.nmax; ++n1)
                if (modes[n1].m == st.second)
                    groups[g].push_back(n1);
        }
        if (static_cast<int>(groups.size()) != contract_groups(plan.nmax))
            throw Error(kInternal, "contraction modes do not pack into their blocks");
        for (size_t g = 0; g < groups.size(); ++g)
            for (size_t i = 0; i < groups[g].size(); ++i)
                row_mode[g * kContractModes + i] = groups[g][i];
    }
    plan.crow.assign(2 * static_cast<size_t>(crows) + 1, 0);
    for (int k = 0; k < crows; ++k) {
        const int n1 = row_mode[k];
        plan.crow[crows + 1 + k] = n1;
        if (n1 < 0) {
            plan.crow[k + 1] = static_cast<int>(plan.cidx.size());
            continue;
        }
        const int m1 = modes[n1].m;
        for (int n2 = 0; n2 < plan.nmax; ++n2) {
            const int dm = modes[n2].m - m1;
            const int dmi = dm + 2 * lmax;
            for (int p = 0; p < P; ++p) {
                const cplx a = ab5[(static_cast<size_t>(n2) * plan.nmax + n1) * P + p];
                if (a == cplx(0.0, 0.0))
                    continue;
                const auto &cl = cols[dmi];
                const int c = static_cast<int>(std::find(cl.begin(), cl.end(), n2) - cl.begin());
                auto it = slot.find({dmi, p, c});
                if (it == slot.end() || c >= static_cast<int>(cl.size()))
                    throw Error(kInternal, "ab5 entry outside the pair plan");
                plan.cidx.push_back(it->second);
                plan.cval.push_back(a.real());
                plan.cval.push_back(a.imag());
                ++plan.nnz_ab5;
            }
        }
        plan.crow[k + 1] = static_cast<int>(plan.cidx.size());
    }
    // every Q slot the contraction reads is computed by every bin list
    for (int form = 0; form < 2; ++form)
        for (int c = 0; c < kHalfBinCaps; ++c) {
            std::vector<char> cov(static_cast<size_t>(plan.nq), 0);
            for (const HalfBin &b : plan.half_bins[form][c])
                for (int w = 0; w < kHalfBinWaves; ++w)
                    for (int o = 0; o < 2; ++o)
                        if (b.wd[w] >= 0 && b.wip[w][o] >= 0)
                            for (int k = b.wtb[w]; k < b.wte[w]; ++k)
                                cov[b.wq[w][o] + k] = 1;
            for (int ci : plan.cidx)
                if (!cov[ci / 16])
                    throw Error(kInternal, "fp16 GEMM bins miss a Q slot of the contraction");
        }

    // Normalised Legendre recurrences (legendre_normalized_trigon.py:59-91):
    //   P_m^m = sqrt((2m+1)/2/(2m)!) (2m-1)!! sin^m
    //   P_{l+1}^m = alpha(l,m) cos P_l^m - beta(l,m) P_{l-1}^m
    plan.leg_alpha.assign(P * P, 0.0);
    plan.leg_beta.assign(P * P, 0.0);
    plan.leg_diag.assign(P, 0.0);
    for (int m = 0; m < P; ++m) {
        long double f2m = 1.0L, dfac = 1.0L;
        for (int i = 1; i <= 2 * m; ++i)
            f2m *= i;
        for (int i = 1; i < 2 * m; i += 2)
            dfac *= i;
        plan.leg_diag[m] = static_cast<double>(std::sqrt((2.0L * m + 1.0L) / 2.0L / f2m) * dfac);
        for (int l = m; l < P; ++l) {
            plan.leg_alpha[l * P + m] =
                std::sqrt(static_cast<double>((2 * l + 1) * (2 * l + 3)) / ((l + 1 - m) * (l + 1 + m)));
            if (l > m)
                plan.leg_beta[l * P + m] = std::sqrt(static_cast<double>((2 * l + 3) * (l - m) * (l + m)) /
                                                     ((2 * l - 1) * (l + 1 - m) * (l + 1 + m)));
        }
    }
    return plan;
}

namespace {

template <typename T>
T *upload_vec(const std::vector<T> &v, hipStream_t s) {
    void *d = nullptr;
    YH_CHECK(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty())
        YH_CHECK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return static_cast<T *>(d);
}

}  // namespace

DevicePlan upload_device_plan(const CouplingPlan &pl, hipStream_t s) {
    DevicePlan dp{};
    dp.lmax = pl.lmax;
    dp.nmax = pl.nmax;
    dp.maxcols = pl.maxcols;
    dp.nq = pl.nq;
    dp.njobs = static_cast<int>(pl.jobs.size());
    dp.stage_elems = pl.stage_elems;
    dp.jobs = upload_vec(pl.jobs, s);
    for (int g = 0; g < kTableGroupings; ++g) {
        dp.tjobs[g] = upload_vec(pl.table_jobs[g], s);
        dp.ntjobs[g] = static_cast<int>(pl.table_jobs[g].size());
    }
    dp.tjobs_w64 = upload_vec(pl.table_jobs_w64, s);
    dp.ntjobs_w64 = static_cast<int>(pl.table_jobs_w64.size());
    dp.colmap = upload_vec(pl.colmap, s);
    dp.crow = upload_vec(pl.crow, s);
    dp.cidx = upload_vec(pl.cidx, s);
    std::vector<double2> cval(pl.cval.size() / 2);
    for (size_t i = 0; i < cval.size(); ++i)
        cval[i] = make_double2(pl.cval[2 * i], pl.cval[2 * i + 1]);
    dp.cval = upload_vec(cval, s);
    dp.la = upload_vec(pl.leg_alpha, s);
    dp.lb = upload_vec(pl.leg_beta, s);
    dp.ld = upload_vec(pl.leg_diag, s);
    const int L = pl.lmax;
    std::vector<int> dmoff(4 * L + 1);
    for (int d = 0, acc = 0; d < 4 * L + 1; ++d) {
        dmoff[d] = acc;
        acc += 2 * L + 1 - std::abs(d - 2 * L);
    }
    dp.dmoff = upload_vec(dmoff, s);
    dp.order_tiles = upload_vec(pl.order_tiles, s);
    for (int f = 0; f < 2; ++f)
        for (int c = 0; c < kHalfBinCaps; ++c) {
            dp.hbins[f][c] = upload_vec(pl.half_bins[f][c], s);
            dp.nhbins[f][c] = static_cast<int>(pl.half_bins[f][c].size());
            dp.hbin_tiles[f][c] = pl.half_bin_tiles[f][c];
        }
    return dp;
}

void free_device_plan(DevicePlan &dp) {
    void *bufs[] = {dp.jobs,         dp.tjobs[0],     dp.tjobs[1],     dp.tjobs[2],     dp.tjobs_w64,
                    dp.colmap,       dp.crow,         dp.cidx,         dp.cval,         dp.la,
                    dp.lb,           dp.ld,           dp.dmoff,        dp.order_tiles,  dp.hbins[0][0],
                    dp.hbins[0][1],  dp.hbins[0][2],  dp.hbins[1][0],  dp.hbins[1][1],  dp.hbins[1][2]};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    dp = DevicePlan{};
}

}  // namespace yasfhip
