// tree.cpp -- octree, i-groups and interaction lists for the soft-force path.
//
// This stands in for FDPS's TreeForForceLong<ForceSoft, EPISoft, EPJSoft>::
// QuadrupoleWithSymmetrySearch (src/petar.hpp:619-623), which PeTar drives
// with theta, n_leaf_limit and n_group_limit (src/petar.hpp:2741) and whose
// multi-walk driver hands the lists to the GPU functor
// (src/petar.hpp:894-899).  FDPS is not in the reference repository, so the
// semantics are the documented ones of a Barnes-Hut group walk:
//   - particles are sorted along a 63-bit Morton curve of the bounding cube;
//   - a cell with more than n_leaf_limit particles is split into octants;
//   - i-groups are the largest cells with at most n_group_limit particles;
//   - a cell is opened for a group when the minimum distance from the
//     group's bounding box to the cell's centre of mass is below
//     L_cell/theta, or when the cell's box and the group's box, either one
//     widened by the members' r_search, overlap (symmetric search: every j
//     within max(r_search_i, r_search_j) of an i ends up as an EPJ); cells
//     that contain the group are always opened;
//   - opened leaves contribute their particles as EPJ entries, closed cells
//     contribute their quadrupole moment (mass, centre of mass, raw second
//     moment sum m (x-c)(x-c)^T, as SPJQuadrupoleInAndOut) as one SPJ entry.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numeric>
#include <parallel/algorithm>
#include <string>
#include <vector>

#include "../../../include/petar_host.h"

namespace pth {
int host_fail(int code, const std::string &msg);
}
using pth::host_fail;

namespace {

struct Node {
    long begin, end;
    int level;
    int n_child;
    long first_child;
};

inline uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

constexpr int MAX_LEVEL = 21;

}  // namespace

struct pth_tree {
    long n = 0;
    int n_leaf_limit = 20, n_group_limit = 512, nthreads = 0;
    std::vector<long> order;
    std::vector<double> pos, mass, rs;
    std::vector<uint64_t> key;
    double root_min[3] = {0, 0, 0}, root_len = 1.0;
    std::vector<Node> nodes;
    std::vector<long> level_start;  // nodes of level l are [level_start[l], level_start[l+1])
    std::vector<double> nmass, ncom, nquad, box_in, box_out;
    std::vector<long> groups;  // node index of each group, Morton order
    std::vector<int> n_epj, n_spj;
    std::vector<long> epj_off, spj_off;
    std::vector<int> id_epj, id_spj;
    bool walked = false;
};

namespace {

void build_nodes(pth_tree &t)
{
    t.nodes.clear();
    t.level_start.clear();
    t.nodes.push_back(Node{0, t.n, 0, 0, -1});
    t.level_start.push_back(0);
    long lb = 0;
    while (lb < (long)t.nodes.size()) {
        long le = (long)t.nodes.size();
        t.level_start.push_back(le);
        const long nl = le - lb;
        std::vector<std::array<long, 9>> bnd(nl);
        std::vector<long> cnt(nl, 0);
#pragma omp parallel for schedule(dynamic, 64)
        for (long k = 0; k < nl; k++) {
            const Node &nd = t.nodes[lb + k];
            const long m = nd.end - nd.begin;
            if (m <= t.n_leaf_limit || nd.level >= MAX_LEVEL) continue;
            const int shift = 3 * (MAX_LEVEL - 1 - nd.level);
            const uint64_t *kb = t.key.data() + nd.begin, *ke = t.key.data() + nd.end;
            std::array<long, 9> &b = bnd[k];
            b[0] = nd.begin;
            b[8] = nd.end;
            for (int d = 1; d < 8; d++) {
                const uint64_t *p = std::partition_point(
                    kb, ke, [&](uint64_t x) { return (int)((x >> shift) & 7u) < d; });
                b[d] = nd.begin + (p - kb);
            }
            long c = 0;
            for (int d = 0; d < 8; d++) c += (b[d + 1] > b[d]);
            cnt[k] = c;
        }
        std::vector<long> off(nl + 1, 0);
        for (long k = 0; k < nl; k++) off[k + 1] = off[k] + cnt[k];
        t.nodes.resize(le + off[nl]);
        for (long k = 0; k < nl; k++) {
            Node &nd = t.nodes[lb + k];
            if (cnt[k] == 0) continue;
            nd.first_child = le + off[k];
            nd.n_child = (int)cnt[k];
            long c = le + off[k];
            for (int d = 0; d < 8; d++)
                if (bnd[k][d + 1] > bnd[k][d])
                    t.nodes[c++] = Node{bnd[k][d], bnd[k][d + 1], nd.level + 1, 0, -1};
        }
        lb = le;
    }
}

void build_moments(pth_tree &t)
{
    const long nn = (long)t.nodes.size();
    t.nmass.assign(nn, 0.0);
    t.ncom.assign(3 * nn, 0.0);
    t.nquad.assign(6 * nn, 0.0);
    t.box_in.assign(6 * nn, 0.0);
    t.box_out.assign(6 * nn, 0.0);
    const int nlev = (int)t.level_start.size() - 1;
    for (int l = nlev - 1; l >= 0; l--) {
        const long a = t.level_start[l], b = t.level_start[l + 1];
#pragma omp parallel for schedule(dynamic, 64)
        for (long k = a; k < b; k++) {
            const Node &nd = t.nodes[k];
            double M = 0, c[3] = {0, 0, 0}, q[6] = {0, 0, 0, 0, 0, 0};
            double bi[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
            double bo[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
            if (nd.n_child == 0) {
                for (long i = nd.begin; i < nd.end; i++) {
                    M += t.mass[i];
                    for (int d = 0; d < 3; d++) {
                        const double x = t.pos[3 * i + d];
                        c[d] += t.mass[i] * x;
                        bi[d] = std::min(bi[d], x);
                        bi[3 + d] = std::max(bi[3 + d], x);
                        bo[d] = std::min(bo[d], x - t.rs[i]);
                        bo[3 + d] = std::max(bo[3 + d], x + t.rs[i]);
                    }
                }
                if (M > 0) {
                    for (int d = 0; d < 3; d++) c[d] /= M;
                } else {
                    for (int d = 0; d < 3; d++) c[d] = 0.5 * (bi[d] + bi[3 + d]);
                }
                for (long i = nd.begin; i < nd.end; i++) {
                    const double x = t.pos[3 * i] - c[0], y = t.pos[3 * i + 1] - c[1],
                                 z = t.pos[3 * i + 2] - c[2], m = t.mass[i];
                    q[0] += m * x * x;
                    q[1] += m * y * y;
                    q[2] += m * z * z;
                    q[3] += m * x * y;
                    q[4] += m * x * z;
                    q[5] += m * y * z;
                }
            } else {
                for (int ch = 0; ch < nd.n_child; ch++) {
                    const long cc = nd.first_child + ch;
                    M += t.nmass[cc];
                    for (int d = 0; d < 3; d++) {
                        c[d] += t.nmass[cc] * t.ncom[3 * cc + d];
                        bi[d] = std::min(bi[d], t.box_in[6 * cc + d]);
                        bi[3 + d] = std::max(bi[3 + d], t.box_in[6 * cc + 3 + d]);
                        bo[d] = std::min(bo[d], t.box_out[6 * cc + d]);
                        bo[3 + d] = std::max(bo[3 + d], t.box_out[6 * cc + 3 + d]);
                    }
                }
                if (M > 0) {
                    for (int d = 0; d < 3; d++) c[d] /= M;
                } else {
                    for (int d = 0; d < 3; d++) c[d] = 0.5 * (bi[d] + bi[3 + d]);
                }
                for (int ch = 0; ch < nd.n_child; ch++) {
                    const long cc = nd.first_child + ch;
                    const double m = t.nmass[cc];
                    const double x = t.ncom[3 * cc] - c[0], y = t.ncom[3 * cc + 1] - c[1],
                                 z = t.ncom[3 * cc + 2] - c[2];
                    const double *qc = &t.nquad[6 * cc];
                    q[0] += qc[0] + m * x * x;
                    q[1] += qc[1] + m * y * y;
                    q[2] += qc[2] + m * z * z;
                    q[3] += qc[3] + m * x * y;
                    q[4] += qc[4] + m * x * z;
                    q[5] += qc[5] + m * y * z;
                }
            }
            t.nmass[k] = M;
            for (int d = 0; d < 3; d++) t.ncom[3 * k + d] = c[d];
            for (int d = 0; d < 6; d++) {
                t.nquad[6 * k + d] = q[d];
                t.box_in[6 * k + d] = bi[d];
                t.box_out[6 * k + d] = bo[d];
            }
        }
    }
}

inline bool overlap(const double *a, const double *b)
{
    for (int d = 0; d < 3; d++)
        if (a[d] > b[3 + d] || b[d] > a[3 + d]) return false;
    return true;
}

inline double mindist2(const double *box, const double *p)
{
    double s = 0;
    for (int d = 0; d < 3; d++) {
        double e = 0;
        if (p[d] < box[d]) e = box[d] - p[d];
        else if (p[d] > box[3 + d]) e = p[d] - box[3 + d];
        s += e * e;
    }
    return s;
}

}  // namespace

extern "C" int pth_tree_build(pth_tree **tree, long n, const double *pos, const double *mass,
                              const double *r_search, int n_leaf_limit, int n_group_limit,
                              int nthreads)
{
    if (!tree) return host_fail(-1, "pth_tree_build: null output pointer");
    *tree = nullptr;
    if (n <= 0) return host_fail(-1, "pth_tree_build: n must be positive");
    if (n > (long)std::numeric_limits<int>::max())
        return host_fail(-1, "pth_tree_build: n exceeds the int32 index range of the lists");
    if (n_leaf_limit < 1 || n_group_limit < 1)
        return host_fail(-1, "pth_tree_build: limits must be >= 1");
    pth_tree *t = new pth_tree;
    t->n = n;
    t->n_leaf_limit = n_leaf_limit;
    t->n_group_limit = n_group_limit;
    t->nthreads = nthreads;
    int saved = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);

    // bounding cube
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (long i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], pos[3 * i + d]);
            hi[d] = std::max(hi[d], pos[3 * i + d]);
        }
    double len = 0;
    for (int d = 0; d < 3; d++) len = std::max(len, hi[d] - lo[d]);
    len = (len > 0) ? len * (1.0 + 1e-9) : 1.0;
    for (int d = 0; d < 3; d++) t->root_min[d] = 0.5 * (lo[d] + hi[d]) - 0.5 * len;
    t->root_len = len;

    // Morton keys and sort
    std::vector<std::pair<uint64_t, long>> kv(n);
    const double scale = (double)(1u << MAX_LEVEL) / len;
#pragma omp parallel for
    for (long i = 0; i < n; i++) {
        uint64_t c[3];
        for (int d = 0; d < 3; d++) {
            double f = (pos[3 * i + d] - t->root_min[d]) * scale;
            long v = (long)std::floor(f);
            v = std::min(std::max(v, 0L), (long)(1u << MAX_LEVEL) - 1);
            c[d] = (uint64_t)v;
        }
        kv[i] = {spread3(c[0]) << 2 | spread3(c[1]) << 1 | spread3(c[2]), i};
    }
    __gnu_parallel::sort(kv.begin(), kv.end());
    t->order.resize(n);
    t->key.resize(n);
    t->pos.resize(3 * n);
    t->mass.resize(n);
    t->rs.resize(n);
#pragma omp parallel for
    for (long k = 0; k < n; k++) {
        const long i = kv[k].second;
        t->order[k] = i;
        t->key[k] = kv[k].first;
        for (int d = 0; d < 3; d++) t->pos[3 * k + d] = pos[3 * i + d];
        t->mass[k] = mass[i];
        t->rs[k] = r_search ? r_search[i] : 0.0;
    }
    build_nodes(*t);
    build_moments(*t);
    if (nthreads > 0) omp_set_num_threads(saved);
    *tree = t;
    return 0;
}

extern "C" int pth_tree_walk(pth_tree *t, double theta)
{
    if (!t) return host_fail(-1, "pth_tree_walk: null tree");
    int saved = omp_get_max_threads();
    if (t->nthreads > 0) omp_set_num_threads(t->nthreads);
    // i-groups: largest cells with at most n_group_limit particles, Morton order
    t->groups.clear();
    std::vector<long> st{0};
    while (!st.empty()) {
        long c = st.back();
        st.pop_back();
        const Node &nd = t->nodes[c];
        if (nd.end - nd.begin <= t->n_group_limit || nd.n_child == 0) {
            t->groups.push_back(c);
        } else {
            for (int ch = nd.n_child - 1; ch >= 0; ch--) st.push_back(nd.first_child + ch);
        }
    }
    const long ng = (long)t->groups.size();
    std::vector<std::vector<int>> le(ng), ls(ng);
    const double inv_theta2 = (theta > 0) ? 1.0 / (theta * theta) : 0.0;
#pragma omp parallel
    {
        std::vector<long> stack;
        stack.reserve(1024);
#pragma omp for schedule(dynamic, 4)
        for (long g = 0; g < ng; g++) {
            const long gn = t->groups[g];
            const long gb = t->nodes[gn].begin, ge = t->nodes[gn].end;
            const double *bin = &t->box_in[6 * gn];
            const double *bout = &t->box_out[6 * gn];
            std::vector<int> &E = le[g];
            std::vector<int> &S = ls[g];
            stack.clear();
            stack.push_back(0);
            while (!stack.empty()) {
                const long c = stack.back();
                stack.pop_back();
                const Node &nd = t->nodes[c];
                bool open;
                if (nd.begin <= gb && ge <= nd.end) open = true;
                else if (theta <= 0) open = true;
                else {
                    const double L = t->root_len / (double)(1L << nd.level);
                    open = mindist2(bin, &t->ncom[3 * c]) < L * L * inv_theta2 ||
                           overlap(bout, &t->box_in[6 * c]) || overlap(bin, &t->box_out[6 * c]);
                }
                if (open) {
                    if (nd.n_child == 0) {
                        for (long i = nd.begin; i < nd.end; i++) E.push_back((int)i);
                    } else {
                        for (int ch = nd.n_child - 1; ch >= 0; ch--) stack.push_back(nd.first_child + ch);
                    }
                } else {
                    S.push_back((int)c);
                }
            }
        }
    }
    t->n_epj.resize(ng);
    t->n_spj.resize(ng);
    t->epj_off.assign(ng + 1, 0);
    t->spj_off.assign(ng + 1, 0);
    for (long g = 0; g < ng; g++) {
        t->n_epj[g] = (int)le[g].size();
        t->n_spj[g] = (int)ls[g].size();
        t->epj_off[g + 1] = t->epj_off[g] + t->n_epj[g];
        t->spj_off[g + 1] = t->spj_off[g] + t->n_spj[g];
    }
    t->id_epj.resize(t->epj_off[ng]);
    t->id_spj.resize(t->spj_off[ng]);
#pragma omp parallel for schedule(dynamic, 16)
    for (long g = 0; g < ng; g++) {
        std::copy(le[g].begin(), le[g].end(), t->id_epj.begin() + t->epj_off[g]);
        std::copy(ls[g].begin(), ls[g].end(), t->id_spj.begin() + t->spj_off[g]);
        std::vector<int>().swap(le[g]);
        std::vector<int>().swap(ls[g]);
    }
    t->walked = true;
    if (t->nthreads > 0) omp_set_num_threads(saved);
    return 0;
}

extern "C" int pth_tree_sizes(const pth_tree *t, long *n_node, long *n_group, long *n_epj_list,
                              long *n_spj_list)
{
    if (!t) return host_fail(-1, "pth_tree_sizes: null tree");
    if (n_node) *n_node = (long)t->nodes.size();
    if (n_group) *n_group = t->walked ? (long)t->groups.size() : 0;
    if (n_epj_list) *n_epj_list = t->walked ? t->epj_off.back() : 0;
    if (n_spj_list) *n_spj_list = t->walked ? t->spj_off.back() : 0;
    return 0;
}

extern "C" int pth_tree_order(const pth_tree *t, long *order)
{
    if (!t || !order) return host_fail(-1, "pth_tree_order: null argument");
    std::copy(t->order.begin(), t->order.end(), order);
    return 0;
}

extern "C" int pth_tree_nodes(const pth_tree *t, double *mass, double *com, double *quad,
                              long *begin, long *end, int *level)
{
    if (!t) return host_fail(-1, "pth_tree_nodes: null tree");
    const long nn = (long)t->nodes.size();
    if (mass) std::copy(t->nmass.begin(), t->nmass.end(), mass);
    if (com) std::copy(t->ncom.begin(), t->ncom.end(), com);
    if (quad) std::copy(t->nquad.begin(), t->nquad.end(), quad);
    for (long k = 0; k < nn; k++) {
        if (begin) begin[k] = t->nodes[k].begin;
        if (end) end[k] = t->nodes[k].end;
        if (level) level[k] = t->nodes[k].level;
    }
    return 0;
}

extern "C" int pth_tree_groups(const pth_tree *t, long *epi_begin, int *n_epi, int *n_epj,
                               int *n_spj)
{
    if (!t || !t->walked) return host_fail(-1, "pth_tree_groups: tree not walked");
    const long ng = (long)t->groups.size();
    for (long g = 0; g < ng; g++) {
        const Node &nd = t->nodes[t->groups[g]];
        if (epi_begin) epi_begin[g] = nd.begin;
        if (n_epi) n_epi[g] = (int)(nd.end - nd.begin);
        if (n_epj) n_epj[g] = t->n_epj[g];
        if (n_spj) n_spj[g] = t->n_spj[g];
    }
    return 0;
}

extern "C" int pth_tree_lists(const pth_tree *t, int *id_epj, int *id_spj)
{
    if (!t || !t->walked) return host_fail(-1, "pth_tree_lists: tree not walked");
    if (id_epj) std::copy(t->id_epj.begin(), t->id_epj.end(), id_epj);
    if (id_spj) std::copy(t->id_spj.begin(), t->id_spj.end(), id_spj);
    return 0;
}

extern "C" void pth_tree_free(pth_tree *t) { delete t; }
