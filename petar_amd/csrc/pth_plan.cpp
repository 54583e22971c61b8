// pth_plan.cpp -- work decomposition of a batch of walks (see pth_plan.h).
//
// The fused form cuts every walk into 64-i tiles and, when the walk's list
// is long, its staged list into chunks of at most chunk_j cost units (EPJ
// blocks first, then SPJ records).  Chunk boundaries depend only on the
// walk's own sizes, so a walk gives bit-identical results whatever batch it
// is part of.  Items are then dealt to the 8 XCDs as 8 contiguous runs of
// equal total cost (neighbouring items share walks, hence staged lists, so
// each list is fetched into one XCD's L2), most expensive first inside each
// run, and written in the order the hardware hands workgroups to XCDs
// (round robin), padding short runs with empty items.
#include "pth_plan.h"

#include <algorithm>
#include <climits>

#include "../../include/petar_hip.h"

namespace pth {

bool make_plan(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj, int exec_mode,
               int chunk_j, Plan &P)
{
    P = Plan();
    P.walks.resize(n_walk);
    P.es_off.assign(n_walk + 1, 0);
    P.ss_off.assign(n_walk + 1, 0);
    long long io = 0, eo = 0, so = 0;
    for (int w = 0; w < n_walk; w++) {
        WalkDev &W = P.walks[w];
        W.i_off = (int)io;
        W.n_i = n_epi[w];
        W.e_off = (int)eo;
        W.n_e = n_epj[w];
        W.s_off = (int)so;
        W.n_s = n_spj[w];
        W.pad[0] = W.pad[1] = 0;
        io += n_epi[w];
        eo += n_epj[w];
        so += n_spj[w];
        P.es_off[w + 1] = P.es_off[w] + ((long long)n_epj[w] + 7) / 8 * 8;
        P.ss_off[w + 1] = P.ss_off[w] + n_spj[w];
        P.n_epep += (long long)n_epi[w] * n_epj[w];
        P.n_epsp += (long long)n_epi[w] * n_spj[w];
    }
    P.ni_tot = io;
    P.n_e_slots = P.es_off[n_walk];
    P.n_s_rec = P.ss_off[n_walk];
    if (io > INT_MAX || eo > INT_MAX || so > INT_MAX || P.n_e_slots / 8 > INT_MAX ||
        P.n_s_rec > INT_MAX)
        return false;
    if (exec_mode == PTH_EXEC_PLAIN) return true;

    const long long cj = (chunk_j > 0) ? chunk_j : DEFAULT_CHUNK_J;
    struct Tmp {
        Item it;
        long long cost;
    };
    std::vector<Tmp> seq;
    for (int w = 0; w < n_walk; w++) {
        const WalkDev &W = P.walks[w];
        if (W.n_i <= 0) continue;
        const long long nb = (P.es_off[w + 1] - P.es_off[w]) / 8;
        const long long units = nb * 8 + (long long)SPJ_COST * W.n_s;
        long long nch = (units + cj - 1) / cj;
        if (nch < 1) nch = 1;
        // chunk k covers cost units [k*units/nch, (k+1)*units/nch); EPJ
        // boundaries are rounded to whole blocks of 8
        std::vector<long long> eb(nch + 1), sb(nch + 1);
        for (long long k = 0; k <= nch; k++) {
            const long long u = (k == nch) ? units : (k * units) / nch;
            long long e = std::min(u, nb * 8);
            e = (e + 7) / 8;  // blocks
            const long long su = std::max(0LL, u - nb * 8);
            eb[k] = std::min(e, nb);
            sb[k] = std::min((su + SPJ_COST - 1) / SPJ_COST, (long long)W.n_s);
        }
        eb[nch] = nb;
        sb[nch] = W.n_s;
        const int n_tiles = (W.n_i + 63) / 64;
        for (int t = 0; t < n_tiles; t++) {
            const int part_first = P.n_part;
            for (long long k = 0; k < nch; k++) {
                Item it;
                it.walk = w;
                it.i_begin = W.i_off + 64 * t;
                it.n_i = std::min(64, W.n_i - 64 * t);
                it.e_blk_begin = (int)(P.es_off[w] / 8 + eb[k]);
                it.e_blk_end = (int)(P.es_off[w] / 8 + eb[k + 1]);
                it.s_begin = (int)(P.ss_off[w] + sb[k]);
                it.s_end = (int)(P.ss_off[w] + sb[k + 1]);
                it.part = (nch > 1) ? P.n_part++ : -1;
                const long long c = 8LL * (it.e_blk_end - it.e_blk_begin) +
                                    (long long)SPJ_COST * (it.s_end - it.s_begin) + 8;
                seq.push_back({it, c});
            }
            if (nch > 1) P.ritems.push_back(RItem{W.i_off + 64 * t, std::min(64, W.n_i - 64 * t),
                                                  part_first, (int)nch});
        }
    }
    P.n_real_items = (int)seq.size();
    // 8 contiguous runs of equal cost
    long long total = 0;
    for (auto &s : seq) total += s.cost;
    std::vector<std::vector<Item>> run(XCDS);
    {
        long long acc = 0;
        int x = 0;
        size_t k = 0;
        std::vector<Tmp> cur;
        auto flush = [&](int xx) {
            std::stable_sort(cur.begin(), cur.end(),
                             [](const Tmp &a, const Tmp &b) { return a.cost > b.cost; });
            for (auto &c : cur) run[xx].push_back(c.it);
            cur.clear();
        };
        for (; k < seq.size(); k++) {
            // item goes to run x while the running cost is below (x+1)/8 of the total
            while (x < XCDS - 1 && acc >= (total * (x + 1)) / XCDS) {
                flush(x);
                x++;
            }
            cur.push_back(seq[k]);
            acc += seq[k].cost;
        }
        flush(x);
    }
    size_t max_blocks = 0;
    for (auto &r : run) max_blocks = std::max(max_blocks, (r.size() + ITEMS_PER_BLOCK - 1) / ITEMS_PER_BLOCK);
    Item dummy;
    dummy.walk = -1;
    dummy.i_begin = dummy.n_i = dummy.e_blk_begin = dummy.e_blk_end = dummy.s_begin = dummy.s_end = 0;
    dummy.part = -1;
    P.items.assign(max_blocks * XCDS * ITEMS_PER_BLOCK, dummy);
    for (int x = 0; x < XCDS; x++)
        for (size_t k = 0; k < run[x].size(); k++) {
            const size_t r = k / ITEMS_PER_BLOCK, q = k % ITEMS_PER_BLOCK;
            P.items[(r * XCDS + x) * ITEMS_PER_BLOCK + q] = run[x][k];
        }
    return true;
}

}  // namespace pth
