// pth_plan.cpp -- work decomposition of a batch of walks (see pth_plan.h).
//
// The fused form cuts every walk into tiles of 64, 128 or 256 i-particles
// (1, 2 or 4 per lane), choosing for each walk the tiling of least issue
// cost: more i per lane share each j record's per-wave scalar load and loop
// overhead, fewer leave fewer lanes idle on small walks.  When the walk's list is long it
// is cut into chunks of at most chunk_j list entries (an SPJ counts twice),
// EPJ blocks first, then SPJ records.  Tiling and chunk boundaries depend
// only on the walk's own sizes, so a walk gives bit-identical results
// whatever batch it is part of.  Items are then dealt to the 8 XCDs as 8
// contiguous runs of equal total cost (neighbouring items share walks,
// hence staged lists, so each list is fetched into one XCD's L2), most
// expensive first inside each run, and written in the order the hardware
// hands workgroups to XCDs (round robin), padding short runs with empty
// items.  With several segments (contiguous walk ranges of equal cost, the
// first of half that, pipelined over several streams) the dealing is done
// per segment.
#include "pth_plan.h"

#include <algorithm>
#include <omp.h>
#include <array>
#include <climits>
#include <cstdint>
#include <cstring>

#include "../../include/petar_hip.h"

namespace pth {

namespace {

struct Tile {
    int begin, n, ipl;
};

struct Tmp {
    Item it;
    double cost;
};

// Working memory of make_plan, kept per thread between calls: a plan of a
// large evaluation touches tens of MB, and fresh pages cost more than the
// planning itself.
// The items of one segment and ipl class and their 8 runs (one per XCD).
struct ClassRuns {
    std::vector<uint32_t> sub;  // items of the class, walk order
    size_t run_begin[XCDS + 1];
    size_t max_blocks = 0;
};

struct Scratch {
    std::vector<Tmp> seq;
    std::vector<size_t> walk_item_begin;
    std::vector<int> walk_ritem_begin;
    std::vector<long long> walk_part_begin;
    std::vector<double> wcost;
    std::vector<uint16_t> bucket;
    std::vector<std::array<ClassRuns, PLAN_CLASSES>> cr;
};

// Runs f(begin, end) over [0, n) on up to 16 host threads, at least `grain`
// indices each (fewer, shorter loops stay on the calling thread).
template <class F> void parallel_for(int n, int grain, F f)
{
    // OpenMP's thread pool: a plan makes several of these calls, and
    // starting threads for each one cost about a millisecond apiece
    const int t = std::min({omp_get_max_threads(), 16, std::max(1, n / std::max(grain, 1))});
    if (t <= 1) {
        f(0, n);
        return;
    }
#pragma omp parallel for num_threads(t) schedule(static, 1)
    for (int k = 0; k < t; k++) f((int)((long long)n * k / t), (int)((long long)n * (k + 1) / t));
}

constexpr int PLAN_WALKS_PER_THREAD = 512;
constexpr int PLAN_ITEMS_PER_THREAD = 32768;

// Least-cost cover of n i-particles by tiles of 256 (ipl 4), 192 (ipl 3),
// 128 (ipl 2) and 64 (ipl 1) lanes x i, for a walk with n_e EPJ and n_s
// SPJ; ties go to the cover with fewer 1-i tiles (the least efficient
// class: most idle lanes, least work per scalar load).
void tile_walk(int n, long long n_e, long long n_s, int max_ipl, std::vector<Tile> &out)
{
    out.clear();
    if (n <= 0) return;
    auto cost = [&](int ipl) { return 64.0 * ipl * (epj_unit_cost(ipl) * n_e + spj_unit_cost(ipl) * n_s); };
    // the largest allowed class any number of times, each smaller one (but
    // the 1-i remainder) at most 3 times
    const int amax = (max_ipl >= 4) ? (n + 255) / 256 : 0;
    const int dmax = (max_ipl >= 4) ? 3 : (max_ipl == 3) ? (n + 191) / 192 : 0;
    const int bmax = (max_ipl >= 3) ? 3 : (max_ipl == 2) ? (n + 127) / 128 : 0;
    double best = -1;
    int ba = 0, bd = 0, bb = 0, bc = 0;
    for (int a = 0; a <= amax; a++)
        for (int d = 0; d <= dmax; d++)
            for (int b = 0; b <= bmax; b++) {
                const long long covered = 256LL * a + 192LL * d + 128LL * b;
                const int c = covered >= n ? 0 : (int)((n - covered + 63) / 64);
                const double t = a * cost(4) + d * cost(3) + b * cost(2) + c * cost(1);
                if (best < 0 || t < best - 1e-9 || (t < best + 1e-9 && c < bc)) {
                    best = t;
                    ba = a;
                    bd = d;
                    bb = b;
                    bc = c;
                }
            }
    int pos = 0;
    auto emit = [&](int count, int ipl) {
        for (int k = 0; k < count && pos < n; k++) {
            const int m = std::min(64 * ipl, n - pos);
            out.push_back(Tile{pos, m, ipl});
            pos += m;
        }
    };
    emit(ba, 4);
    emit(bd, 3);
    emit(bb, 2);
    emit(bc, 1);
}

}  // namespace

bool make_plan(int n_walk, const int *n_epi, const int *n_epj, const int *n_spj, int exec_mode,
               int chunk_j, int max_ipl, Plan &P, int n_seg, bool mono, long long min_seg_pairs)
{
    // the item array keeps its size from the previous plan (every slot is
    // written below, so growing it is the only initialisation it needs)
    std::vector<Item> kept;
    kept.swap(P.items);
    P.reset();
    P.items.swap(kept);
    P.mono = mono;
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
        P.es_off[w + 1] = P.es_off[w] + epj_region_slots((long long)n_epj[w] + (mono ? n_spj[w] : 0));
        P.ss_off[w + 1] = P.ss_off[w] + (mono ? 0 : n_spj[w]);
        P.n_epep += (long long)n_epi[w] * n_epj[w];
        P.n_epsp += (long long)n_epi[w] * n_spj[w];
        if (io > INT_MAX || eo > INT_MAX || so > INT_MAX) return false;
    }
    P.ni_tot = io;
    P.n_e_slots = P.es_off[n_walk];
    P.n_s_rec = P.ss_off[n_walk];
    if (P.n_e_slots / 8 > INT_MAX || P.n_s_rec > INT_MAX) return false;
    if (exec_mode == PTH_EXEC_PLAIN) {
        PlanSeg g;
        g.walk_end = n_walk;
        P.segs.push_back(g);
        return true;
    }
    if (max_ipl < 1 || max_ipl > 4) max_ipl = 4;

    const long long cj = (chunk_j > 0) ? chunk_j : DEFAULT_CHUNK_J;
    static thread_local Scratch S;
    // chunking of walk w: nch chunks, EPJ block / SPJ record bounds
    auto chunks = [&](int w, std::vector<long long> &eb, std::vector<long long> &sb) -> long long {
        const long long nb = (P.es_off[w + 1] - P.es_off[w]) / 8;
        const long long n_s = P.ss_off[w + 1] - P.ss_off[w];  // staged SPJ records
        const long long units = nb * 8 + 2LL * n_s;
        long long nch = (units + cj - 1) / cj;
        if (nch < 1) nch = 1;
        // chunk k covers list units [k*units/nch, (k+1)*units/nch); EPJ
        // boundaries are rounded to whole blocks of 8
        eb.resize(nch + 1);
        sb.resize(nch + 1);
        for (long long k = 0; k <= nch; k++) {
            const long long u = (k == nch) ? units : (k * units) / nch;
            eb[k] = std::min((std::min(u, nb * 8) + 7) / 8, nb);
            sb[k] = std::min((std::max(0LL, u - nb * 8) + 1) / 2, n_s);
        }
        eb[nch] = nb;
        sb[nch] = n_s;
        return nch;
    };
    // 1. per walk: items (tiles x chunks), reduction items and partial slots
    std::vector<size_t> &walk_item_begin = S.walk_item_begin;
    std::vector<int> &walk_ritem_begin = S.walk_ritem_begin;
    std::vector<long long> &walk_part_begin = S.walk_part_begin;
    walk_item_begin.assign(n_walk + 1, 0);
    walk_ritem_begin.assign(n_walk + 1, 0);
    walk_part_begin.assign(n_walk + 1, 0);
    parallel_for(n_walk, PLAN_WALKS_PER_THREAD, [&](int wb, int we) {
        std::vector<long long> eb, sb;
        std::vector<Tile> tiles;
        for (int w = wb; w < we; w++) {
            const WalkDev &W = P.walks[w];
            if (W.n_i <= 0) continue;
            const long long nch = chunks(w, eb, sb);
            tile_walk(W.n_i, mono ? (long long)W.n_e + W.n_s : W.n_e, mono ? 0 : W.n_s, max_ipl, tiles);
            long long ipl_sum = 0;
            for (const Tile &t : tiles) ipl_sum += t.ipl;
            walk_item_begin[w + 1] = tiles.size() * (size_t)nch;
            walk_ritem_begin[w + 1] = nch > 1 ? (int)tiles.size() : 0;
            walk_part_begin[w + 1] = nch > 1 ? nch * ipl_sum : 0;
        }
    });
    for (int w = 0; w < n_walk; w++) {
        walk_item_begin[w + 1] += walk_item_begin[w];
        walk_ritem_begin[w + 1] += walk_ritem_begin[w];
        walk_part_begin[w + 1] += walk_part_begin[w];
    }
    if (walk_part_begin[n_walk] > INT_MAX) return false;
    std::vector<Tmp> &seq = S.seq;
    seq.resize(walk_item_begin[n_walk]);
    // wcost[w + 1]: cost of walk w's items, summed into prefix form below
    std::vector<double> &wcost = S.wcost;
    wcost.assign(n_walk + 1, 0.0);
    P.ritems.resize(walk_ritem_begin[n_walk]);
    P.n_part = (int)walk_part_begin[n_walk];
    parallel_for(n_walk, PLAN_WALKS_PER_THREAD, [&](int wb, int we) {
        std::vector<long long> eb, sb;
        std::vector<Tile> tiles;
        for (int w = wb; w < we; w++) {
            const WalkDev &W = P.walks[w];
            if (W.n_i <= 0) continue;
            const long long nch = chunks(w, eb, sb);
            tile_walk(W.n_i, mono ? (long long)W.n_e + W.n_s : W.n_e, mono ? 0 : W.n_s, max_ipl, tiles);
            size_t o = walk_item_begin[w];
            int r_o = walk_ritem_begin[w];
            int part = (int)walk_part_begin[w];
            double wc = 0.0;
            for (const Tile &t : tiles) {
                const int part_first = part;
                for (long long k = 0; k < nch; k++) {
                    Item it;
                    it.walk = w;
                    it.i_begin = W.i_off + t.begin;
                    it.n_i_ipl = item_pack(t.n, t.ipl);
                    it.e_blk_begin = (int)(P.es_off[w] / 8 + eb[k]);
                    it.e_blk_end = (int)(P.es_off[w] / 8 + eb[k + 1]);
                    it.s_begin = (int)(P.ss_off[w] + sb[k]);
                    it.s_end = (int)(P.ss_off[w] + sb[k + 1]);
                    it.part = -1;
                    if (nch > 1) {
                        it.part = part;
                        part += t.ipl;
                    }
                    const double c = 64.0 * t.ipl *
                                     (epj_unit_cost(t.ipl) * 8.0 * (it.e_blk_end - it.e_blk_begin) +
                                      spj_unit_cost(t.ipl) * (it.s_end - it.s_begin) + 50.0);
                    seq[o++] = Tmp{it, c};
                    wc += c;
                }
                if (nch > 1) {
                    RItem r;
                    r.out = W.i_off + t.begin;
                    r.n_i = t.n;
                    r.part_first = part_first;
                    r.n_part = (int)nch;
                    r.ipl = t.ipl;
                    r.pad[0] = r.pad[1] = r.pad[2] = 0;
                    P.ritems[r_o++] = r;
                }
            }
            wcost[w + 1] = wc;
        }
    });
    P.n_real_items = (int)seq.size();
    // 2. segments: contiguous walk ranges of about equal item cost (the
    //    first one half of it)
    if (n_seg < 1) n_seg = 1;
    if (n_seg > MAX_SEGMENTS) n_seg = MAX_SEGMENTS;
    if (n_seg > n_walk) n_seg = std::max(n_walk, 1);
    // a pipeline pays for its launches and cross-stream events, so each
    // segment gets at least min_seg_pairs interactions
    if (min_seg_pairs > 0)
        n_seg = (int)std::min<long long>(n_seg, std::max<long long>(1, (P.n_epep + P.n_epsp) / min_seg_pairs));
    for (int w = 0; w < n_walk; w++) wcost[w + 1] += wcost[w];
    // the first segment gets half the cost of the others: its staging is the
    // one part of a pipelined evaluation that nothing overlaps
    std::vector<int> cut(n_seg + 1, 0);
    cut[n_seg] = n_walk;
    for (int g = 1; g < n_seg; g++) {
        const double target = wcost[n_walk] * (g - 0.5) / (n_seg - 0.5);
        cut[g] = (int)(std::lower_bound(wcost.begin(), wcost.end(), target) - wcost.begin());
        cut[g] = std::min(std::max(cut[g], cut[g - 1]), n_walk);
    }
    // 3. cost buckets for the most-expensive-first order inside a run: a
    //    counting sort on the cost's binary exponent and top 4 mantissa bits
    //    (16 steps per octave), stable inside a bucket
    constexpr int NBUCKET = 1024;
    std::vector<uint16_t> &bucket = S.bucket;
    bucket.resize(seq.size());
    parallel_for((int)std::min<size_t>(seq.size(), INT_MAX), PLAN_ITEMS_PER_THREAD, [&](int kb, int ke) {
        for (int k = kb; k < ke; k++) {
            const double c = std::max(seq[k].cost, 1.0);  // >= 1: biased exponent >= 1023
            uint64_t bits;
            std::memcpy(&bits, &c, 8);
            const int q = (int)((bits >> 48) - (1023ull << 4));  // exponent * 16 + top mantissa bits
            bucket[k] = (uint16_t)std::min(NBUCKET - 1, q);
        }
    });
    // 4. per segment and ipl class (one launch each; each kernel has its
    //    own register budget): 8 contiguous runs of equal cost, one per XCD.
    //    First the runs and sizes of every segment, then the items, both in
    //    parallel over segments.
    static const int class_ipl[PLAN_CLASSES] = {1, 2, 3, 4};
    std::vector<std::array<ClassRuns, PLAN_CLASSES>> &cr = S.cr;
    cr.resize(n_seg);
    parallel_for(n_seg, 1, [&](int gb, int ge) {
        for (int g = gb; g < ge; g++)
            for (int cls = 0; cls < PLAN_CLASSES; cls++) {
                ClassRuns &R = cr[g][cls];
                R.sub.clear();
                double total = 0;
                for (size_t k = walk_item_begin[cut[g]]; k < walk_item_begin[cut[g + 1]]; k++)
                    if (item_ipl(seq[k].it.n_i_ipl) == class_ipl[cls]) {
                        R.sub.push_back((uint32_t)k);
                        total += seq[k].cost;
                    }
                // an item goes to run x while the running cost is below
                // (x+1)/8 of the total
                double acc = 0;
                int x = 0;
                R.run_begin[0] = 0;
                for (size_t k = 0; k < R.sub.size(); k++) {
                    while (x < XCDS - 1 && acc >= total * (x + 1) / XCDS) R.run_begin[++x] = k;
                    acc += seq[R.sub[k]].cost;
                }
                while (x < XCDS - 1) R.run_begin[++x] = R.sub.size();
                R.run_begin[XCDS] = R.sub.size();
                R.max_blocks = 0;
                for (int xx = 0; xx < XCDS; xx++)
                    R.max_blocks = std::max(R.max_blocks, (R.run_begin[xx + 1] - R.run_begin[xx] + ITEMS_PER_BLOCK - 1) /
                                                              ITEMS_PER_BLOCK);
            }
    });
    size_t cursor = 0;
    P.segs.resize(n_seg);
    for (int g = 0; g < n_seg; g++) {
        PlanSeg &sg = P.segs[g];
        sg.walk_begin = cut[g];
        sg.walk_end = cut[g + 1];
        sg.ritem_begin = walk_ritem_begin[cut[g]];
        sg.ritem_count = walk_ritem_begin[cut[g + 1]] - sg.ritem_begin;
        for (int cls = 0; cls < PLAN_CLASSES; cls++) {
            sg.seg_begin[cls] = (int)cursor;
            const size_t sz = cr[g][cls].sub.empty() ? 0 : cr[g][cls].max_blocks * XCDS * ITEMS_PER_BLOCK;
            sg.seg_count[cls] = (int)sz;
            cursor += sz;
        }
    }
    if (cursor > (size_t)INT_MAX) return false;
    Item dummy;
    dummy.walk = -1;
    dummy.i_begin = dummy.n_i_ipl = dummy.e_blk_begin = dummy.e_blk_end = dummy.s_begin = dummy.s_end = 0;
    dummy.part = -1;
    P.items.resize(cursor);
    parallel_for(n_seg, 1, [&](int gb, int ge) {
        std::vector<uint32_t> cnt(NBUCKET + 1), sorted;
        for (int g = gb; g < ge; g++)
            for (int cls = 0; cls < PLAN_CLASSES; cls++) {
                const ClassRuns &R = cr[g][cls];
                if (R.sub.empty()) continue;
                const size_t base = P.segs[g].seg_begin[cls];
                for (int xx = 0; xx < XCDS; xx++) {
                    const size_t a = R.run_begin[xx], b = R.run_begin[xx + 1];
                    // descending bucket order, stable
                    std::fill(cnt.begin(), cnt.end(), 0u);
                    for (size_t k = a; k < b; k++) cnt[NBUCKET - 1 - bucket[R.sub[k]] + 1]++;
                    for (int q = 0; q < NBUCKET; q++) cnt[q + 1] += cnt[q];
                    sorted.resize(b - a);
                    for (size_t k = a; k < b; k++) sorted[cnt[NBUCKET - 1 - bucket[R.sub[k]]]++] = R.sub[k];
                    // the run's slots: its items, then empty items up to the
                    // longest run's length
                    const size_t n_slot = (size_t)R.max_blocks * ITEMS_PER_BLOCK;
                    for (size_t k = 0; k < n_slot; k++) {
                        const size_t r = k / ITEMS_PER_BLOCK, q = k % ITEMS_PER_BLOCK;
                        P.items[base + (r * XCDS + xx) * ITEMS_PER_BLOCK + q] =
                            k < sorted.size() ? seq[sorted[k]].it : dummy;
                    }
                }
            }
    });
    for (int cls = 0; cls < PLAN_CLASSES; cls++) {  // whole-plan view (one segment)
        P.seg_begin[cls] = P.segs[0].seg_begin[cls];
        P.seg_count[cls] = P.segs[0].seg_count[cls];
    }
    return true;
}

}  // namespace pth
